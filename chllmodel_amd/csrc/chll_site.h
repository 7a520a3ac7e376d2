// chll_site.h -- per-site work of the sweep, shared by the single-lattice
// kernels (chll_kernels.hip) and the batched replica kernels
// (chll_ensemble.hip): one director trial, one chirality trial, one site pair's
// observables, and the block reductions.
#pragma once

#include "chll_kernels.h"

namespace chll {

// Device-scope atomic add through a global-address-space pointer: a
// global_atomic even where the compiler cannot tell that the pointer
// addresses global memory (see fst below).
__device__ inline void atomic_add_global(uint32_t *p, uint32_t v)
{
    __hip_atomic_fetch_add((__attribute__((address_space(1))) uint32_t *)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Block-wide sum of one 0/1 flag per thread, added to slot blk % kSlots.
template <int BS>
__device__ inline void block_count(bool flag, uint32_t *slots, uint32_t blk)
{
    __shared__ uint32_t wsum[BS / 64];
    const uint64_t bal = __ballot(flag);
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (lane == 0) wsum[wid] = (uint32_t)__popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < BS / 64; ++w) t += wsum[w];
        if (t) atomic_add_global(&slots[(blk & (kSlots - 1)) * kSlotStride], t);
    }
}

// Element q of a field.  OFF32: addressed by a 32-bit byte offset from the
// field's base, which compiles to one global load or store with a scalar base
// and a 32-bit vector offset instead of 64-bit address arithmetic per field;
// valid while a field spans less than 4 GiB (field_off32).
template <bool OFF32, class T>
__device__ inline T &fld(T *base, uint32_t q)
{
    if (OFF32) return *(T *)((char *)base + q * (uint32_t)sizeof(T));
    return base[q];
}

// Block-wide sums of two per-wave counts w0, w1 (the same in every lane of a
// wave), added to slot blk % kSlots of slots0 and slots1 (thread 0 and thread
// 1 add).
template <int BS>
__device__ inline void block_sum2(uint32_t w0, uint32_t w1, uint32_t *slots0, uint32_t *slots1, uint32_t blk)
{
    __shared__ uint32_t wsum[2][BS / 64];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (lane == 0) {
        wsum[0][wid] = w0;
        wsum[1][wid] = w1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < BS / 64; ++w) t += wsum[threadIdx.x][w];
        uint32_t *slots = threadIdx.x ? slots1 : slots0;
        if (t) atomic_add_global(&slots[(blk & (kSlots - 1)) * kSlotStride], t);
    }
}

// Block-wide sums of two 0/1 flags per thread.
template <int BS>
__device__ inline void block_count2(bool f0, bool f1, uint32_t *slots0, uint32_t *slots1, uint32_t blk)
{
    block_sum2<BS>((uint32_t)__popcll(__ballot(f0)), (uint32_t)__popcll(__ballot(f1)), slots0, slots1, blk);
}

// Store v to element q of a field through a global-address-space pointer, so
// that the store is a global_store even where the compiler cannot tell that
// the field pointer addresses global memory (pointers read from device
// memory, as the replica descriptors are); it would otherwise emit a flat
// store, which also counts against lgkmcnt.
template <bool OFF32, class T>
__device__ inline void fst(T *base, uint32_t q, T v)
{
    T *p = &fld<OFF32>(base, q);
    *(__attribute__((address_space(1))) T *)p = v;
}

template <bool OFF32>
__device__ inline void load_nbrs(const Fields &F, const uint32_t nb[6], Nbr &N)
{
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const uint32_t q = nb[b];
        N.x[b] = fld<OFF32>(F.nx, q);
        N.y[b] = fld<OFF32>(F.ny, q);
        N.z[b] = fld<OFF32>(F.nz, q);
        N.s[b] = fld<OFF32>(F.s, q);
    }
}

// A trial's inputs: the site's director and spin, its six neighbours (order
// +x, -x, +y, -y, +z, -z, absent ones zero) and their presence mask.
struct SiteIn {
    Nbr N;
    double nx, ny, nz;
    int sg;
};

// Keeps the compiler from reusing loaded neighbour values across it, so an
// accessor that reads LDS reloads them instead of holding them in registers
// through the trial move.
__device__ inline void reload_barrier() { __asm__ volatile("" ::: "memory"); }

// Makes x, y, z depend on v without changing them, so that the arithmetic
// that follows on x, y, z starts after v is complete: a rarely taken branch
// that evaluates two site energies then holds one energy's temporaries at a
// time instead of interleaving both.
__device__ inline void after(double v, double &x, double &y, double &z)
{
    __asm__ volatile("" : "+v"(x), "+v"(y), "+v"(z) : "v"(v));
}

// Director trial (chll/chll_mod.f90:130-200) with neighbour accessor N, the
// site's director (nx, ny, nz) and spin sg, and its uniforms u_phi (cone
// angle) and u_acc (acceptance); returns whether it is accepted and the trial
// director in t[3].
template <int ARITH, class NB>
__device__ inline bool dir_decide(const PhaseArgs &A, const NB &N, double nx, double ny, double nz, int sg, double d,
                                  double u_phi, double u_acc, double t[3])
{
    if (ARITH == ARITH_COMPAT) {
        // the reference's binary32-rounded energies around the binary64 cone
        // trial (trial_fp64): its trial director agrees with the reference's
        // operation order (chll/chll_mod.f90:149-183, IEEE divisions and
        // square roots) to a few ulp, which the energies' binary32 roundings
        // absorb; the reference's order cost a fifth of the director kernel.
        // The binary64 difference decides unless the energies' roundings
        // could change the outcome (compat_screen); those trials evaluate
        // the rounded energies.
        double c2o, c2n, mo, mn;
        const double ao = cqf_sum_mag(N, sg, nx, ny, nz, c2o, mo);
        trial_fp64(nx, ny, nz, d, u_phi, t[0], t[1], t[2]);
        reload_barrier();
        const double an = cqf_sum_mag(N, sg, t[0], t[1], t[2], c2n, mn);
        const double dE = (c2o - c2n) - A.KK * (an - ao);
        const double M = (12.0 + (c2o + c2n)) + fabs(A.KK) * (mo + mn) + 2.0 * fabs(A.g * (double)sg);
        const int sc = compat_screen(dE, M, A.screen, A.kbt, A.inv_kbt, u_acc);
        if (sc >= 0) return sc != 0;
        reload_barrier();
        const double eold = energy_compat(N, nx, ny, nz, sg, A.g, A.KK);
        after(eold, t[0], t[1], t[2]);
        const double enew = energy_compat(N, t[0], t[1], t[2], sg, A.g, A.KK);
        return (enew < eold) || accept_exp(div_by_recip(-(enew - eold), A.kbt, A.inv_kbt), u_acc);
    } else {
        double c2o, c2n;
        const double ao = cqf_sum(N, sg, nx, ny, nz, c2o);
        trial_fp64(nx, ny, nz, d, u_phi, t[0], t[1], t[2]);
        reload_barrier();
        const double an = cqf_sum(N, sg, t[0], t[1], t[2], c2n);
        const double dE = (c2o - c2n) - A.KK * (an - ao);
        return (dE < 0.0) || accept_exp(-dE * A.inv_kbt, u_acc);
    }
}

// The director trial of packed site m of `colour` in sweep A.istep with its
// own uniforms (dir_decide).
template <int ARITH, int RNG, class NB>
__device__ inline bool dir_compute(const PhaseArgs &A, int colour, uint32_t m, const NB &N, double nx, double ny,
                                   double nz, int sg, double d, double t[3])
{
    double u_phi, u_acc;
    if (RNG == RNG_PHILOX) {
        const U4 w = site_block(A.seed, A.run_index, A.replica, m, A.istep, colour);
        u_phi = site_u_phi(w);
        u_acc = site_u_dir(w);
    } else {
        u_phi = A.u_phi[m];
        u_acc = A.u_acc[m];
    }
    return dir_decide<ARITH>(A, N, nx, ny, nz, sg, d, u_phi, u_acc, t);
}

// Chirality trial (chll/chll_mod.f90:214-243).
template <int ARITH, int RNG, class NB>
__device__ inline bool spin_compute(const PhaseArgs &A, int colour, uint32_t m, const NB &N, double nx, double ny,
                                    double nz, int sg)
{
    double u_acc;
    if (RNG == RNG_PHILOX) {
        u_acc = site_u_spin(site_block(A.seed, A.run_index, A.replica, m, spin_sweep(colour, A.istep), colour));
    } else {
        u_acc = A.u_acc[m];
    }
    if (ARITH == ARITH_COMPAT) {
        double csq, w;
        const double a = cq_sum_mag(N, sg, nx, ny, nz, csq, w);
        const double dE = (double)sg * fma(A.KK, a, 2.0 * A.g);
        const double M = fma(2.0, csq, 12.0) + fabs(A.KK) * w + 2.0 * fabs(A.g * (double)sg);
        const int sc = compat_screen(dE, M, A.screen, A.kbt, A.inv_kbt, u_acc);
        if (sc >= 0) return sc != 0;
        reload_barrier();
        const double eold = energy_compat(N, nx, ny, nz, sg, A.g, A.KK);
        double px = nx, py = ny, pz = nz;
        after(eold, px, py, pz);
        const double enew = energy_compat(N, px, py, pz, -sg, A.g, A.KK);
        return (enew < eold) || accept_exp(div_by_recip(-(enew - eold), A.kbt, A.inv_kbt), u_acc);
    } else {
        const double a = cq_sum(N, nx, ny, nz);
        const double dE = (double)sg * fma(A.KK, a, 2.0 * A.g);
        return (dE < 0.0) || accept_exp(-dE * A.inv_kbt, u_acc);
    }
}

// Gather a site's inputs straight from the colour-split fields.
template <int BC, bool OFF32>
__device__ inline uint32_t load_site(const Geom &G, const Fields &F, int colour, uint32_t m, SiteIn &S)
{
    const Site st = decode(G, m, colour);
    uint32_t nb[6];
    const uint32_t own = colour ? G.nsub + m : m;
    const uint32_t oth = colour ? 0u : G.nsub;
    S.N.mask = neighbours<BC>(G, m, st, oth, nb);
    load_nbrs<OFF32>(F, nb, S.N);
    S.nx = fld<OFF32>(F.nx, own); S.ny = fld<OFF32>(F.ny, own); S.nz = fld<OFF32>(F.nz, own);
    S.sg = fld<OFF32>(F.s, own);
    return own;
}

// Trials on loaded inputs S of packed site m of `colour` whose field index is
// `own`.  Accepted moves are stored at once, or (DEFER) the site's director
// and spin after its trials go to *o, for a caller that stores two sites with
// 16-byte words (store_pair).
struct SiteOut {
    double nx, ny, nz;
    int s;
};

template <bool OFF32, bool DEFER>
__device__ inline void put_dir(const Fields &F, uint32_t own, bool acc, const double t[3], const SiteIn &S,
                               SiteOut *o)
{
    if constexpr (DEFER) {
        o->nx = acc ? t[0] : S.nx;
        o->ny = acc ? t[1] : S.ny;
        o->nz = acc ? t[2] : S.nz;
    } else if (acc) {
        fst<OFF32>(F.nx, own, t[0]);
        fst<OFF32>(F.ny, own, t[1]);
        fst<OFF32>(F.nz, own, t[2]);
    }
}

template <bool OFF32, bool DEFER>
__device__ inline void put_spin(const Fields &F, uint32_t own, bool flip, int sg_new, SiteOut *o)
{
    if constexpr (DEFER) o->s = sg_new;
    else if (flip) fst<OFF32>(F.s, own, (int8_t)sg_new);
}

// One director trial; returns whether it is accepted.
// d: the cone step, read from *A.d once per thread by the caller (a second
// read after a trial would be a dependent load in the middle of the work).
template <int ARITH, int RNG, bool OFF32, bool DEFER = false>
__device__ inline bool dir_apply(const Fields &F, int colour, const PhaseArgs &A, double d, uint32_t m, uint32_t own,
                                 const SiteIn &S, SiteOut *o = nullptr)
{
    double t[3];
    const bool acc = dir_compute<ARITH, RNG>(A, colour, m, S.N, S.nx, S.ny, S.nz, S.sg, d, t);
    put_dir<OFF32, DEFER>(F, own, acc, t, S, o);
    if constexpr (DEFER) o->s = S.sg;
    return acc;
}

// One chirality trial; returns whether the spin flips.
template <int ARITH, int RNG, bool OFF32, bool DEFER = false>
__device__ inline bool spin_apply(const Fields &F, int colour, const PhaseArgs &A, uint32_t m, uint32_t own,
                                  const SiteIn &S, SiteOut *o = nullptr)
{
    const bool acc = spin_compute<ARITH, RNG>(A, colour, m, S.N, S.nx, S.ny, S.nz, S.sg);
    put_spin<OFF32, DEFER>(F, own, acc, acc ? -S.sg : S.sg, o);
    if constexpr (DEFER) {
        o->nx = S.nx;
        o->ny = S.ny;
        o->nz = S.nz;
    }
    return acc;
}

// Director trial, then chirality trial; FP64 arithmetic only.  The reference
// runs director colour 0, director colour 1, chirality colour 0, chirality
// colour 1 (chll/chll_mod.f90:70-81).  In FP64 arithmetic the chirality energy
// change sigma (K sum_b sgn_b c_b q_b + 2 g) reads the directors and the
// site's own spin but no neighbour spin, so once both colours' directors are
// final a site's chirality trial has the same outcome before or after the
// other colour's chirality trials.  Run right after this site's director trial
// it reuses the neighbour directors in registers and saves one pass over the
// lattice.  Returns the director decision; `flip` gets the chirality decision.
// FP64 chirality decision (spin_compute) from the site's chirality sum a and
// its uniform u_acc.
__device__ inline bool spin_decide_fp64(const PhaseArgs &A, double inv_kbt, double a, int sg, double u_acc)
{
    const double dE = (double)sg * fma(A.KK, a, 2.0 * A.g);
    return (dE < 0.0) || accept_exp(-dE * inv_kbt, u_acc);
}

// The director trial (dir_compute, FP64) then the chirality trial on the
// accepted director (spin_compute, FP64); the bond passes of the director
// trial also give the chirality sums of the old and the trial director.
template <int RNG, bool OFF32, bool DEFER = false>
__device__ inline bool dir_spin_apply(const Fields &F, int colour, const PhaseArgs &A, double d, uint32_t m,
                                      uint32_t own, const SiteIn &S, bool &flip, SiteOut *o = nullptr)
{
    // for colour 1 (the fused launch) one block serves both trials
    double u_phi, u_acc, u_spin;
    if (RNG == RNG_PHILOX) {
        const U4 w = site_block(A.seed, A.run_index, A.replica, m, A.istep, colour);
        u_phi = site_u_phi(w);
        u_acc = site_u_dir(w);
        u_spin = site_u_spin(colour ? w : site_block(A.seed, A.run_index, A.replica, m, spin_sweep(0, A.istep), 0));
    } else {
        u_phi = A.u_phi[m];
        u_acc = A.u_acc[m];
        u_spin = A.u_spin[m];
    }
    double t[3], c2o, c2n, cqo, cqn;
    const double ao = cqf_cq_sum(S.N, S.sg, S.nx, S.ny, S.nz, c2o, cqo);
    trial_fp64(S.nx, S.ny, S.nz, d, u_phi, t[0], t[1], t[2]);
    reload_barrier();
    const double an = cqf_cq_sum(S.N, S.sg, t[0], t[1], t[2], c2n, cqn);
    const double dE = (c2o - c2n) - A.KK * (an - ao);
    const bool acc = (dE < 0.0) || accept_exp(-dE * A.inv_kbt, u_acc);
    put_dir<OFF32, DEFER>(F, own, acc, t, S, o);
    flip = spin_decide_fp64(A, A.inv_kbt, acc ? cqn : cqo, S.sg, u_spin);
    put_spin<OFF32, DEFER>(F, own, flip, flip ? -S.sg : S.sg, o);
    return acc;
}

// Chirality trial of sweep A.istep_spin, then director trial of sweep A.istep,
// at a colour-0 site; FP64 arithmetic, Philox draws.  After sweep t's colour-1
// director and chirality trials every input of a colour-0 site for its
// chirality trial in sweep t (own director, colour-1 directors) and for its
// director trial in sweep t + 1 (own director and spin after the flip,
// colour-1 directors and spins) is final, and by the argument of
// dir_spin_apply no colour-0 chirality trial reads another colour-0 site.  The
// pair therefore gives the sequential outcome; the cone step of sweep t + 1 is
// fixed before this launch, since it depends on sweep t's director accepts
// only (chll/chll_mod.f90:83-90).  Returns the director decision.
template <bool OFF32, bool DEFER = false>
__device__ inline bool spin_dir_apply(const Fields &F, const PhaseArgs &A, double d, uint32_t m, uint32_t own,
                                      const SiteIn &S, bool &flip, SiteOut *o = nullptr)
{
    // one pass over the old director's bonds: the chirality sum, then the
    // director's old sum from the same bond products with the spin the
    // chirality trial leaves
    double c2o, cqb[6];
    const double cqo = cq_products(S.N, S.nx, S.ny, S.nz, c2o, cqb);
    // one block: the chirality trial of sweep A.istep_spin and the director
    // trial of sweep A.istep = A.istep_spin + 1 both draw from block A.istep
    // (spin_sweep); every caller sets A.istep that way
    const U4 w = site_block(A.seed, A.run_index, A.replica, m, A.istep, 0);
    flip = spin_decide_fp64(A, A.inv_kbt_spin, cqo, S.sg, site_u_spin(w));
    const int sg = flip ? -S.sg : S.sg;
    // director trial of sweep A.istep (dir_compute, FP64)
    const double u_phi = site_u_phi(w), u_acc = site_u_dir(w);
    double t[3], c2n;
    const double ao = weight_cq(S.N, sg, cqb);
    trial_fp64(S.nx, S.ny, S.nz, d, u_phi, t[0], t[1], t[2]);
    reload_barrier();
    const double an = cqf_sum(S.N, sg, t[0], t[1], t[2], c2n);
    const double dE = (c2o - c2n) - A.KK * (an - ao);
    const bool acc = (dE < 0.0) || accept_exp(-dE * A.inv_kbt, u_acc);
    put_spin<OFF32, DEFER>(F, own, flip, sg, o);
    put_dir<OFF32, DEFER>(F, own, acc, t, S, o);
    return acc;
}

// ---------------------------------------------------------------------------
// Reference arithmetic in fused launches: held chirality trials.
//
// The reference runs director colour 0, director colour 1, chirality colour 0,
// chirality colour 1 (chll/chll_mod.f90:70-81), and the rounded site energies
// of a chirality trial read the neighbours' spins.  The screen
// (compat_screen) decides a chirality trial from the binary64 difference
// sigma (K sum_b sgn_b c_b q_b + 2 g) and a magnitude that reads only |s_b|.
// No flip changes either, so a screened decision of a colour-1 site is the
// same before or after colour 0's chirality trials, and the site's chirality
// trial runs right after its director trial, in one launch as in FP64.  A
// trial the screen leaves open is held: the spin keeps its value, the site
// joins A.held, and k_held decides it once its six colour-0 neighbours'
// chirality outcomes are computed (each a function of the state the colour-1
// launch leaves and of that neighbour's own draw), then flips it, before
// colour 0's launch.  Colour 0's chirality trials thus
// run after colour 1's; the rounded energies of one that the screen leaves
// open read the colour-1 spins before those flips, recovered through
// A.flipbits (a flip negates the spin).

// The screen of a reference-arithmetic chirality trial (spin_compute):
// 1 flip, 0 keep, -1 open.
template <class NB>
__device__ inline int spin_screen_compat(const PhaseArgs &A, double kbt, double inv_kbt, const NB &N, double nx,
                                         double ny, double nz, int sg, double u)
{
    double csq, w;
    const double a = cq_sum_mag(N, sg, nx, ny, nz, csq, w);
    const double dE = (double)sg * fma(A.KK, a, 2.0 * A.g);
    const double M = fma(2.0, csq, 12.0) + fabs(A.KK) * w + 2.0 * fabs(A.g * (double)sg);
    return compat_screen(dE, M, A.screen, kbt, inv_kbt, u);
}

// The reference's chirality decision from its two rounded energies
// (chll/chll_mod.f90:214-243), as spin_compute evaluates them.
template <class NB>
__device__ inline bool spin_exact_compat(const PhaseArgs &A, double kbt, double inv_kbt, const NB &N, double nx,
                                         double ny, double nz, int sg, double u)
{
    const double eold = energy_compat(N, nx, ny, nz, sg, A.g, A.KK);
    double px = nx, py = ny, pz = nz;
    after(eold, px, py, pz);
    const double enew = energy_compat(N, px, py, pz, -sg, A.g, A.KK);
    return (enew < eold) || accept_exp(div_by_recip(-(enew - eold), kbt, inv_kbt), u);
}

// The rounded-energy decision of a chirality trial of colour-0 packed site m
// (reference arithmetic).  With A.flipbits set, colour 1's chirality trials of
// this sweep have run, and the trial reads its colour-1 neighbours' spins from
// before their flips: N's spins are changed to those, and flipped[b] marks
// the ones changed (a caller that runs a director trial on N afterwards
// changes them back).
template <int BC>
__device__ inline bool spin_c0_exact(const Geom &G, const PhaseArgs &A, double kbt, double inv_kbt, uint32_t m,
                                     Nbr &N, double nx, double ny, double nz, int sg, double u, int8_t flipped[6])
{
    if (A.flipbits) {
        // the site's neighbours recomputed here, from an index the compiler
        // cannot relate to the caller's: otherwise it keeps the caller's
        // neighbour indices alive through the trial for this rare branch
        uint32_t mm = m;
        __asm__ volatile("" : "+v"(mm));
        const Site st = decode(G, mm, 0);
        uint32_t nb[6];
        neighbours<BC>(G, mm, st, 0u, nb);  // packed colour-1 indices
        const uint32_t *w32 = (const uint32_t *)A.flipbits;  // bit m % 32 of word m / 32
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            if (((N.mask >> b) & 1u) && ((w32[nb[b] >> 5] >> (nb[b] & 31u)) & 1u)) {
                N.s[b] = -N.s[b];
                flipped[b] = 1;
            }
        }
    }
    return spin_exact_compat(A, kbt, inv_kbt, N, nx, ny, nz, sg, u);
}

// Chirality trial of colour-0 packed site m in reference arithmetic: the
// screen, else spin_c0_exact.
template <int BC>
__device__ inline bool spin_c0_compat(const Geom &G, const PhaseArgs &A, double kbt, double inv_kbt, uint32_t m,
                                      Nbr &N, double nx, double ny, double nz, int sg, double u, int8_t flipped[6])
{
    const int sc = spin_screen_compat(A, kbt, inv_kbt, N, nx, ny, nz, sg, u);
    if (sc >= 0) return sc != 0;
    reload_barrier();
    return spin_c0_exact<BC>(G, A, kbt, inv_kbt, m, N, nx, ny, nz, sg, u, flipped);
}

// Colour-0 chirality trial of sweep A.istep alone (reference arithmetic).
template <int RNG, int BC>
__device__ inline bool spin_c0_apply(const Geom &G, const PhaseArgs &A, uint32_t m, SiteIn &S, SiteOut *o)
{
    const double u = RNG == RNG_PHILOX
                         ? site_u_spin(site_block(A.seed, A.run_index, A.replica, m, spin_sweep(0, A.istep), 0))
                         : A.u_acc[m];
    int8_t fl[6] = {0, 0, 0, 0, 0, 0};
    const bool flip = spin_c0_compat<BC>(G, A, A.kbt, A.inv_kbt, m, S.N, S.nx, S.ny, S.nz, S.sg, u, fl);
    o->nx = S.nx;
    o->ny = S.ny;
    o->nz = S.nz;
    o->s = flip ? -S.sg : S.sg;
    return flip;
}

// Reference-arithmetic director trial (dir_decide) given the old director's
// screen sums: ao (cqf_sum), c2o and mo (its magnitude).
template <class NB>
__device__ inline bool dir_decide_compat(const PhaseArgs &A, const NB &N, double nx, double ny, double nz, int sg,
                                         double d, double u_phi, double u_acc, double ao, double c2o, double mo,
                                         double t[3])
{
    double c2n, mn;
    trial_fp64(nx, ny, nz, d, u_phi, t[0], t[1], t[2]);
    reload_barrier();
    const double an = cqf_sum_mag(N, sg, t[0], t[1], t[2], c2n, mn);
    const double dE = (c2o - c2n) - A.KK * (an - ao);
    const double M = (12.0 + (c2o + c2n)) + fabs(A.KK) * (mo + mn) + 2.0 * fabs(A.g * (double)sg);
    const int sc = compat_screen(dE, M, A.screen, A.kbt, A.inv_kbt, u_acc);
    if (sc >= 0) return sc != 0;
    reload_barrier();
    const double eold = energy_compat(N, nx, ny, nz, sg, A.g, A.KK);
    after(eold, t[0], t[1], t[2]);
    const double enew = energy_compat(N, t[0], t[1], t[2], sg, A.g, A.KK);
    return (enew < eold) || accept_exp(div_by_recip(-(enew - eold), A.kbt, A.inv_kbt), u_acc);
}

// Colour-1 director trial, then its chirality trial on the director it
// leaves (reference arithmetic): `flip` gets a screened flip, `held` an open
// trial (the spin keeps its value).  Returns the director decision.  The
// director trial's two bond passes also give both directors' chirality sums
// for the chirality screen, whose magnitude sum_b |c_b q_b| max(|s|, |s_b|)
// is bounded by max(|s|, max_b |s_b|) sum_b |c_b q_b| (equal for spins +-1;
// a larger bound only holds more trials).  A director trial that takes the
// rounded energies recomputes the chirality sums afterwards instead of
// keeping them through that branch.
template <int RNG>
__device__ inline bool dir_spin_held(const PhaseArgs &A, double d, uint32_t m, const SiteIn &S, bool &flip,
                                     bool &held, SiteOut *o)
{
    double u_phi, u_acc, u_spin;
    if (RNG == RNG_PHILOX) {
        const U4 w = site_block(A.seed, A.run_index, A.replica, m, A.istep, 1);
        u_phi = site_u_phi(w);
        u_acc = site_u_dir(w);
        u_spin = site_u_spin(w);
    } else {
        u_phi = A.u_phi[m];
        u_acc = A.u_acc[m];
        u_spin = A.u_spin[m];
    }
    double c2o, mo, cqo, abo, t[3], c2n, mn, cqn, abn;
    const double ao = cqf_sum_mag_cq(S.N, S.sg, S.nx, S.ny, S.nz, c2o, mo, cqo, abo);
    trial_fp64(S.nx, S.ny, S.nz, d, u_phi, t[0], t[1], t[2]);
    reload_barrier();
    const double an = cqf_sum_mag_cq(S.N, S.sg, t[0], t[1], t[2], c2n, mn, cqn, abn);
    const double dEd = (c2o - c2n) - A.KK * (an - ao);
    const double Md = (12.0 + (c2o + c2n)) + fabs(A.KK) * (mo + mn) + 2.0 * fabs(A.g * (double)S.sg);
    const int scd = compat_screen(dEd, Md, A.screen, A.kbt, A.inv_kbt, u_acc);
    bool acc;
    double a, csq, w;
    if (scd >= 0) {
        acc = scd != 0;
        int smax = abs(S.sg);
#pragma unroll
        for (int b = 0; b < 6; ++b) smax = max(smax, abs(S.N.spin(b)));
        a = acc ? cqn : cqo;
        csq = acc ? c2n : c2o;
        w = (acc ? abn : abo) * (double)smax;
    } else {
        reload_barrier();
        const double eold = energy_compat(S.N, S.nx, S.ny, S.nz, S.sg, A.g, A.KK);
        after(eold, t[0], t[1], t[2]);
        const double enew = energy_compat(S.N, t[0], t[1], t[2], S.sg, A.g, A.KK);
        acc = (enew < eold) || accept_exp(div_by_recip(-(enew - eold), A.kbt, A.inv_kbt), u_acc);
        reload_barrier();
        a = acc ? cq_sum_mag(S.N, S.sg, t[0], t[1], t[2], csq, w) : cq_sum_mag(S.N, S.sg, S.nx, S.ny, S.nz, csq, w);
    }
    o->nx = acc ? t[0] : S.nx;
    o->ny = acc ? t[1] : S.ny;
    o->nz = acc ? t[2] : S.nz;
    // the chirality screen (spin_screen_compat) on the director the trial left
    const double dE = (double)S.sg * fma(A.KK, a, 2.0 * A.g);
    const double M = fma(2.0, csq, 12.0) + fabs(A.KK) * w + 2.0 * fabs(A.g * (double)S.sg);
    const int sc = compat_screen(dE, M, A.screen, A.kbt, A.inv_kbt, u_spin);
    flip = sc == 1;
    held = sc < 0;
    o->s = flip ? -S.sg : S.sg;
    return acc;
}

// Colour-0 chirality trial of sweep A.istep_spin, then the director trial of
// sweep A.istep = A.istep_spin + 1 (reference arithmetic, Philox; both from
// one block, spin_sweep), after colour 1's held trials are decided: the
// director trial reads the colour-1 spins after every flip of sweep
// A.istep_spin, as the reference's next sweep does.  One pass over the old
// director's bonds serves the chirality screen and the director trial's old
// sums.  Returns the director decision.
template <int BC>
__device__ inline bool spin_dir_held(const Geom &G, const PhaseArgs &A, double d, uint32_t m, SiteIn &S, bool &flip,
                                     SiteOut *o)
{
    const U4 w = site_block(A.seed, A.run_index, A.replica, m, A.istep, 0);
    const double u_spin = site_u_spin(w);
    double c2o, cqb[6], wc;
    const double a = cq_products_mag(S.N, S.sg, S.nx, S.ny, S.nz, c2o, cqb, wc);
    {
        // the chirality screen (spin_screen_compat), else the rounded energies
        const double dE = (double)S.sg * fma(A.KK, a, 2.0 * A.g);
        const double M = fma(2.0, c2o, 12.0) + fabs(A.KK) * wc + 2.0 * fabs(A.g * (double)S.sg);
        const int sc = compat_screen(dE, M, A.screen, A.kbt_spin, A.inv_kbt_spin, u_spin);
        if (sc >= 0) {
            flip = sc != 0;
        } else {
            reload_barrier();
            int8_t fl[6] = {0, 0, 0, 0, 0, 0};
            flip = spin_c0_exact<BC>(G, A, A.kbt_spin, A.inv_kbt_spin, m, S.N, S.nx, S.ny, S.nz, S.sg, u_spin, fl);
#pragma unroll
            for (int b = 0; b < 6; ++b)
                if (fl[b]) S.N.s[b] = -S.N.s[b];  // the director trial reads the spins after colour 1's flips
        }
    }
    const int sg = flip ? -S.sg : S.sg;
    double mo, t[3];
    const double ao = weight_cq_mag(S.N, sg, cqb, mo);
    const bool acc = dir_decide_compat(A, S.N, S.nx, S.ny, S.nz, sg, d, site_u_phi(w), site_u_dir(w), ao, c2o, mo, t);
    o->nx = acc ? t[0] : S.nx;
    o->ny = acc ? t[1] : S.ny;
    o->nz = acc ? t[2] : S.nz;
    o->s = sg;
    return acc;
}

// Interleaves the bits of v with zeros (bit b -> bit 2 b).
__device__ inline unsigned long long spread32(uint32_t v)
{
    unsigned long long x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

// After a fused colour-1 launch's trials in reference arithmetic (k_pair,
// k_pair_ens): each wave whose lanes hold consecutive site pairs m0 = 2 (64 w
// + lane) writes the flip bits of its 128 sites (two 64-bit words, site m at
// bit m % 64 of word m / 64), and appends the sites of its held trials to
// A.held.  The ballots take every lane; live_wave: the wave belongs to a live
// block.  The queue holds nsub entries, one per colour-1 site and sweep.
__device__ inline void held_epilogue(const PhaseArgs &A, uint32_t nsub, bool live_wave, uint32_t w, uint32_t m0,
                                     bool f0, bool f1, bool h0, bool h1)
{
    const uint64_t b0 = __ballot(f0), b1 = __ballot(f1), q0 = __ballot(h0), q1 = __ballot(h1);
    const uint32_t lane = threadIdx.x & 63u;
    if (!live_wave) return;
    if (lane == 0u) {
        typedef unsigned long long u2v __attribute__((ext_vector_type(2)));
        typedef __attribute__((address_space(1))) u2v gu2v;
        *(gu2v *)(A.flipbits + 2u * w) = u2v{spread32((uint32_t)b0) | (spread32((uint32_t)b1) << 1),
                                            spread32((uint32_t)(b0 >> 32)) | (spread32((uint32_t)(b1 >> 32)) << 1)};
    }
    if (q0 | q1) {
        uint32_t base = 0;
        if (lane == 0u) base = atomicAdd(A.held, (uint32_t)(__popcll(q0) + __popcll(q1)));
        base = __shfl(base, 0);
        const uint64_t below = (1ull << lane) - 1ull;
        const uint32_t e0 = base + (uint32_t)__popcll(q0 & below);
        const uint32_t e1 = base + (uint32_t)(__popcll(q0) + __popcll(q1 & below));
        if (h0 && e0 < nsub) A.held[kHeldHead + e0] = m0;
        if (h1 && e1 < nsub) A.held[kHeldHead + e1] = m0 + 1u;
    }
}

// Held chirality trials of colour 1 of one lattice, block blk of the nblk
// blocks that serve its queue, kHeldLanes lanes per trial.  Lane b < 6
// computes the chirality outcome of the trial site's colour-0 neighbour b:
// the trial colour 0's next launch runs, on the same inputs and draw (A0,
// spin_c0_compat).  Lane 6 loads the held site's own inputs meanwhile and then
// decides its trial on the colour-0 spins those outcomes leave, as the
// reference's order has them (A1).  The decisions read only the state the
// colour-1 launch left; the last of the nblk blocks to finish applies the
// flips (spin and flip bit) and clears the queue.  The loop over groups of
// trials is uniform, so every lane of a group takes part in the shuffles.
template <int RNG, int BC>
__device__ inline void held_group(const Geom &G, const Fields &F, const PhaseArgs &A0, const PhaseArgs &A1,
                                  uint32_t *held, uint32_t *counts_spin, unsigned long long *total, uint32_t blk,
                                  uint32_t nblk)
{
    constexpr uint32_t kGroupsPerBlock = kHeldBlockThreads / kHeldLanes;
    // the first entry of this group is loaded with the count (in bounds
    // whatever the count; used only below it)
    const uint32_t e_first = blk * kGroupsPerBlock + threadIdx.x / kHeldLanes;
    const uint32_t m_first = held[kHeldHead + min(e_first, G.nsub - 1u)];
    const uint32_t n = min(held[0], G.nsub);
    if (blk == 0 && threadIdx.x == 0 && total) atomicAdd(total, (unsigned long long)n);
    const uint32_t lane = threadIdx.x % kHeldLanes, per = nblk * kGroupsPerBlock;
    uint32_t nflip = 0;
    for (uint32_t e0 = 0; e0 < n; e0 += per) {
        const uint32_t e = e0 + e_first;
        const bool live = e < n;
        const uint32_t m1 = live ? (e0 == 0 ? m_first : held[kHeldHead + e]) : 0u;
        int after_c0 = 0;
        SiteIn S;
        uint32_t mask = 0u;
        if (live) {
            if (lane < 6u) {
                uint32_t nb[6];
                mask = neighbours<BC>(G, m1, decode(G, m1, 1), 0u, nb);  // packed colour-0 indices
                if ((mask >> lane) & 1u) {
                    const uint32_t i = nb[lane];
                    SiteIn Si;
                    load_site<BC, true>(G, F, 0, i, Si);
                    const double u =
                        RNG == RNG_PHILOX
                            ? site_u_spin(site_block(A0.seed, A0.run_index, A0.replica, i, spin_sweep(0, A0.istep), 0))
                            : A0.u_acc[i];
                    int8_t fl[6] = {0, 0, 0, 0, 0, 0};
                    const bool f =
                        spin_c0_compat<BC>(G, A0, A0.kbt, A0.inv_kbt, i, Si.N, Si.nx, Si.ny, Si.nz, Si.sg, u, fl);
                    after_c0 = f ? -Si.sg : Si.sg;
                }
            } else if (lane == 6u) {
                load_site<BC, true>(G, F, 1, m1, S);
                mask = S.N.mask;
            }
        }
        int sb[6];
#pragma unroll
        for (int b = 0; b < 6; ++b) sb[b] = __shfl(after_c0, b, kHeldLanes);
        if (live && lane == 6u) {
#pragma unroll
            for (int b = 0; b < 6; ++b)
                if ((mask >> b) & 1u) S.N.s[b] = sb[b];
            const double u =
                RNG == RNG_PHILOX
                    ? site_u_spin(site_block(A1.seed, A1.run_index, A1.replica, m1, spin_sweep(1, A1.istep), 1))
                    : A1.u_acc[m1];
            const int sc = spin_screen_compat(A1, A1.kbt, A1.inv_kbt, S.N, S.nx, S.ny, S.nz, S.sg, u);
            const bool f = sc >= 0 ? sc != 0 : spin_exact_compat(A1, A1.kbt, A1.inv_kbt, S.N, S.nx, S.ny, S.nz, S.sg, u);
            if (f) {
                held[kHeldHead + e] = m1 | kHeldFlip;
                ++nflip;
            }
        }
    }
    if (nflip) atomic_add_global(counts_spin + (blk % kSlots) * kSlotStride, nflip);
    // The last block applies every decision.  Each wave's fence waits for
    // that wave's own stores, so every thread fences before the barrier:
    // then all of the block's decisions are in device-coherent memory before
    // thread 0 counts the block done, and the last block acquires them with
    // its own fence.
    __shared__ uint32_t is_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) is_last = atomicAdd(held + 1, 1u) == nblk - 1u;
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    for (uint32_t e = threadIdx.x; e < n; e += kHeldBlockThreads) {
        const uint32_t v = __hip_atomic_load(held + kHeldHead + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v & kHeldFlip) {
            const uint32_t m = v & ~kHeldFlip;
            int8_t *p = F.s + G.nsub + m;
            *p = (int8_t)(-*p);
            atomicOr(A1.flipbits + (m >> 6), 1ull << (m & 63u));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        held[0] = 0u;
        held[1] = 0u;
    }
}

// Single sites: load, then apply.
template <int ARITH, int RNG, int BC, bool OFF32 = false>
__device__ inline bool dir_site(const Geom &G, const Fields &F, int colour, const PhaseArgs &A, uint32_t m)
{
    SiteIn S;
    const uint32_t own = load_site<BC, OFF32>(G, F, colour, m, S);
    return dir_apply<ARITH, RNG, OFF32>(F, colour, A, *A.d, m, own, S);
}

template <int ARITH, int RNG, int BC, bool OFF32 = false>
__device__ inline bool spin_site(const Geom &G, const Fields &F, int colour, const PhaseArgs &A, uint32_t m)
{
    SiteIn S;
    const uint32_t own = load_site<BC, OFF32>(G, F, colour, m, S);
    return spin_apply<ARITH, RNG, OFF32>(F, colour, A, m, own, S);
}

template <int RNG, int BC, bool OFF32>
__device__ inline bool dir_spin_site(const Geom &G, const Fields &F, int colour, const PhaseArgs &A, uint32_t m,
                                     bool &flip)
{
    SiteIn S;
    const uint32_t own = load_site<BC, OFF32>(G, F, colour, m, S);
    return dir_spin_apply<RNG, OFF32>(F, colour, A, *A.d, m, own, S, flip);
}

template <int BC, bool OFF32>
__device__ inline bool spin_dir_site(const Geom &G, const Fields &F, const PhaseArgs &A, uint32_t m, bool &flip)
{
    SiteIn S;
    const uint32_t own = load_site<BC, OFF32>(G, F, 0, m, S);
    return spin_dir_apply<OFF32>(F, A, *A.d, m, own, S, flip);
}

// Site pairs.  With half = ni / 2 even, packed sites m0 = 2 P and m0 + 1 of one
// colour lie in one row, and every index below is even: the pair's own fields
// and its y and z neighbours (q, q + 1) are read with 16-byte loads (spins
// with 2-byte loads), an absent y or z bond of both sites with the zero pair.
// Along x the two sites' four bonds touch three other-colour sites q - 1 + p ..
// q + 1 + p (q = oth + m0, p = row parity): the pair (q, q + 1), whose bonds
// always exist, and one more site, the -x neighbour of m0 (p = 0) or the +x
// neighbour of m1 (p = 1), absent, wrapped or at q - 1 / q + 2 as
// neighbours() gives it.  Per pair: 15 field loads of 16 bytes and 3 of 8
// bytes instead of 42 of 8 bytes.
template <bool OFF32>
__device__ inline double2 fld2(const double *base, uint32_t q)
{
    if (OFF32) return *(const double2 *)((const char *)base + q * 8u);
    return *(const double2 *)(base + q);
}

template <bool OFF32>
__device__ inline uint32_t spin2(const int8_t *base, uint32_t q)
{
    if (OFF32) return *(const uint16_t *)((const char *)base + q);
    return *(const uint16_t *)(base + q);
}

__device__ inline int lo_spin(uint32_t v) { return (int)(int8_t)(v & 0xffu); }
__device__ inline int hi_spin(uint32_t v) { return (int)(int8_t)(v >> 8); }

// The kernels that call load_pair give consecutive lanes of a wave consecutive
// pairs (m0 = 2 (tile base + lane)).  The extra x neighbour of a pair is then,
// inside a row, the x pair that the lane beside it loaded: site q - 1 is lane
// l - 1's q + 1 (row parity p = 0), site q + 2 is lane l + 1's q (p = 1).  It
// moves between lanes by DPP; only a pair at a wave's edge whose neighbour
// is in the next wave, or a periodic wrap, loads it (an open row end reads
// zero).  That takes four loads per pair off the memory pipeline.
template <int BC, bool OFF32>
__device__ inline uint32_t load_pair(const Geom &G, const Fields &F, int colour, uint32_t m0, SiteIn &S0, SiteIn &S1)
{
    const Site st0 = decode(G, m0, colour);
    const Site st1{st0.i + 2u, st0.j, st0.k};
    const uint32_t own = (colour ? G.nsub : 0u) + m0;
    const uint32_t oth = colour ? 0u : G.nsub;
    uint32_t n0[6], n1[6];
    S0.N.mask = neighbours<BC>(G, m0, st0, oth, n0);
    S1.N.mask = neighbours<BC>(G, m0 + 1u, st1, oth, n1);
    const uint32_t p = st0.i & 1u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = (st0.i >> 1) >> 1, R = G.half >> 1;  // pair position in its row, pairs per row
    // Every load is issued before any loaded value is used, so that the wave
    // waits for memory once: its own pair, the x pair (q, q + 1), the extra x
    // site where the lane beside it does not hold it, then the y and z pairs.
    const double2 ox = fld2<OFF32>(F.nx, own), oy = fld2<OFF32>(F.ny, own), oz = fld2<OFF32>(F.nz, own);
    const uint32_t os = spin2<OFF32>(F.s, own);
    // x: pair (q, q + 1) and the extra site e
    const uint32_t q = oth + m0;
    const uint32_t e = p ? n1[0] : n0[1];
    const double2 xx = fld2<OFF32>(F.nx, q), xy = fld2<OFF32>(F.ny, q), xz = fld2<OFF32>(F.nz, q);
    const uint32_t xs = spin2<OFF32>(F.s, q);
    const bool from_prev = p == 0u && r != 0u && lane != 0u;
    const bool from_next = p == 1u && r + 1u != R && lane != 63u;
    const bool none = BC == BC_OPEN && (p == 0u ? r == 0u : r + 1u == R);
    double ex = 0.0, ey = 0.0, ez = 0.0;
    int es = 0;
    if (!from_prev && !from_next && !none) {
        ex = fld<OFF32>(F.nx, e); ey = fld<OFF32>(F.ny, e); ez = fld<OFF32>(F.nz, e);
        es = fld<OFF32>(F.s, e);
    }
    // y and z: pairs at site m0's neighbour (its bond and site m1's exist
    // together; absent ones read the zero pair)
    double2 vx[4], vy[4], vz[4];
    uint32_t vs[4];
#pragma unroll
    for (int b = 2; b < 6; ++b) {
        const uint32_t rr = n0[b];
        vx[b - 2] = fld2<OFF32>(F.nx, rr); vy[b - 2] = fld2<OFF32>(F.ny, rr); vz[b - 2] = fld2<OFF32>(F.nz, rr);
        vs[b - 2] = spin2<OFF32>(F.s, rr);
    }
    S0.nx = ox.x; S0.ny = oy.x; S0.nz = oz.x; S0.sg = lo_spin(os);
    S1.nx = ox.y; S1.ny = oy.y; S1.nz = oz.y; S1.sg = hi_spin(os);
    // the extra x neighbour from the adjacent lane: site q - 1 is lane l - 1's
    // q + 1 (row parity 0), site q + 2 is lane l + 1's q (parity 1)
    const double px = lane_prev(xx.y), py = lane_prev(xy.y), pz = lane_prev(xz.y);
    const uint32_t ps = lane_prev(xs);
    const double nxv = lane_next(xx.x), nyv = lane_next(xy.x), nzv = lane_next(xz.x);
    const uint32_t ns = lane_next(xs);
    if (from_prev) {
        ex = px; ey = py; ez = pz; es = hi_spin(ps);
    } else if (from_next) {
        ex = nxv; ey = nyv; ez = nzv; es = lo_spin(ns);
    }
    if (p == 0u) {
        S0.N.x[0] = xx.x; S0.N.y[0] = xy.x; S0.N.z[0] = xz.x; S0.N.s[0] = lo_spin(xs);
        S0.N.x[1] = ex;   S0.N.y[1] = ey;   S0.N.z[1] = ez;   S0.N.s[1] = es;
        S1.N.x[0] = xx.y; S1.N.y[0] = xy.y; S1.N.z[0] = xz.y; S1.N.s[0] = hi_spin(xs);
        S1.N.x[1] = xx.x; S1.N.y[1] = xy.x; S1.N.z[1] = xz.x; S1.N.s[1] = lo_spin(xs);
    } else {
        S0.N.x[0] = xx.y; S0.N.y[0] = xy.y; S0.N.z[0] = xz.y; S0.N.s[0] = hi_spin(xs);
        S0.N.x[1] = xx.x; S0.N.y[1] = xy.x; S0.N.z[1] = xz.x; S0.N.s[1] = lo_spin(xs);
        S1.N.x[0] = ex;   S1.N.y[0] = ey;   S1.N.z[0] = ez;   S1.N.s[0] = es;
        S1.N.x[1] = xx.y; S1.N.y[1] = xy.y; S1.N.z[1] = xz.y; S1.N.s[1] = hi_spin(xs);
    }
#pragma unroll
    for (int b = 2; b < 6; ++b) {
        S0.N.x[b] = vx[b - 2].x; S0.N.y[b] = vy[b - 2].x; S0.N.z[b] = vz[b - 2].x; S0.N.s[b] = lo_spin(vs[b - 2]);
        S1.N.x[b] = vx[b - 2].y; S1.N.y[b] = vy[b - 2].y; S1.N.z[b] = vz[b - 2].y; S1.N.s[b] = hi_spin(vs[b - 2]);
    }
    return own;
}

// Two sites' states after their trials, stored as 16-byte words (directors,
// DIRS) and one 2-byte word (spins, SPINS) at the pair's field index own.  A
// site of the active colour is read by no other thread of the launch, so
// writing back an unchanged value is harmless; every cache line of the active
// colour is written back anyway.  NT: streaming (nontemporal) stores, which
// the replica ensemble takes (its chunks run faster with them; the single
// lattice's launches do not).
template <bool DIRS, bool SPINS, bool NT = false>
__device__ inline void store_pair(const Fields &F, uint32_t own, const SiteOut &O0, const SiteOut &O1)
{
    typedef double d2v __attribute__((ext_vector_type(2)));
    typedef __attribute__((address_space(1))) d2v gd2v;
    typedef __attribute__((address_space(1))) uint16_t gu16;
    auto put = [](auto v, auto *p) {
        if constexpr (NT) __builtin_nontemporal_store(v, p);
        else *p = v;
    };
    if constexpr (DIRS) {
        put(d2v{O0.nx, O1.nx}, (gd2v *)((char *)F.nx + own * 8u));
        put(d2v{O0.ny, O1.ny}, (gd2v *)((char *)F.ny + own * 8u));
        put(d2v{O0.nz, O1.nz}, (gd2v *)((char *)F.nz + own * 8u));
    }
    if constexpr (SPINS)
        put((uint16_t)(((uint32_t)O0.s & 0xffu) | (((uint32_t)O1.s & 0xffu) << 8)), (gu16 *)((char *)F.s + own));
}

// ---------------------------------------------------------------------------
// Observables.  Per site: the reference's reported energy term (periodic i
// and j, open k, +y bond x-component from the diagonal site (i+1, j+1, k),
// chll/chll_mod.f90:427-452), the Hamiltonian's own +x/+y/+z bonds under the
// sweep's boundary condition, the six Q-tensor sums n_a n_b and the spin.
struct MeasureAcc {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long s = 0;
};

// Packed index of site (i, j, k): l >> 1 (any shape with N even).
__device__ inline uint32_t pidx(const Geom &G, uint32_t i, uint32_t j, uint32_t k)
{
    return (i + G.ni * (j + G.nj * k)) >> 1;
}

__device__ inline uint32_t colour_base(const Geom &G, uint32_t i, uint32_t j, uint32_t k)
{
    return ((i + j + k) & 1u) ? G.nsub : 0u;
}

__device__ inline double bond_term(double ax, double ay, double az, double bx, double by, double bz, int axis,
                                   int sa, int sb, double KK)
{
    const double dd = ax * bx + ay * by + az * bz;
    double cr;
    if (axis == 0) cr = ay * bz - az * by;
    else if (axis == 1) cr = az * bx - ax * bz;
    else cr = ax * by - ay * bx;
    const double sf = 0.5 * (double)(sa + sb);
    return (1.0 - dd * dd) - KK * dd * cr * sf;
}

// Both sites of packed index m: linear indices l = 2m and 2m + 1.
template <int BC>
__device__ inline void measure_site(const Geom &G, const Fields &F, double g, double KK, uint32_t m, MeasureAcc &A)
{
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const Site st = decode_l(G, 2u * m + (uint32_t)b);
        const uint32_t i = st.i, j = st.j, k = st.k;
        const uint32_t jp = (j + 1 == G.nj) ? 0u : j + 1;
        const uint32_t ip = (i + 1 == G.ni) ? 0u : i + 1;
        const uint32_t c = colour_base(G, i, j, k) + m;
        const double ax = F.nx[c], ay = F.ny[c], az = F.nz[c];
        const int sa = F.s[c];
        const uint32_t qx = colour_base(G, ip, j, k) + pidx(G, ip, j, k);
        const double bx0 = F.nx[qx], by0 = F.ny[qx], bz0 = F.nz[qx];
        const int sx = F.s[qx];
        const uint32_t qy = colour_base(G, i, jp, k) + pidx(G, i, jp, k);
        const double bx1 = F.nx[qy], by1 = F.ny[qy], bz1 = F.nz[qy];
        const int sy = F.s[qy];
        const uint32_t qd = colour_base(G, ip, jp, k) + pidx(G, ip, jp, k);
        const double dgx = F.nx[qd];
        const double tx = bond_term(ax, ay, az, bx0, by0, bz0, 0, sa, sx, KK);
        // the reference's +y bond: its dot product takes nx of the diagonal site
        double ety;
        {
            const double dd = ax * dgx + ay * by1 + az * bz1;
            const double cr = az * bx1 - ax * bz1;
            const double sf = 0.5 * (double)(sa + sy);
            ety = (1.0 - dd * dd) - KK * dd * cr * sf;
        }
        const double ty = bond_term(ax, ay, az, bx1, by1, bz1, 1, sa, sy, KK);
        double tz = 0.0, tzp = 0.0;
        if (k + 1 < G.nk) {
            const uint32_t qz = colour_base(G, i, j, k + 1) + pidx(G, i, j, k + 1);
            tz = bond_term(ax, ay, az, F.nx[qz], F.ny[qz], F.nz[qz], 2, sa, F.s[qz], KK);
            tzp = tz;
        } else if (BC == BC_PERIODIC) {
            const uint32_t qz = colour_base(G, i, j, 0) + pidx(G, i, j, 0);
            tzp = bond_term(ax, ay, az, F.nx[qz], F.ny[qz], F.nz[qz], 2, sa, F.s[qz], KK);
        }
        const double gs = g * (double)sa;
        A.v[0] += ((tx + ety) + tz) - gs;
        const bool hx = (BC == BC_PERIODIC) || (i + 1 < G.ni);
        const bool hy = (BC == BC_PERIODIC) || (j + 1 < G.nj);
        A.v[1] += ((hx ? tx : 0.0) + (hy ? ty : 0.0) + tzp) - gs;
        A.v[2] += ax * ax;
        A.v[3] += ay * ay;
        A.v[4] += az * az;
        A.v[5] += ax * ay;
        A.v[6] += ax * az;
        A.v[7] += ay * az;
        A.s += sa;
    }
}

// The four sites i = 2 h .. 2 h + 3 of packed indices m, m + 1 of one row, in
// x order, for fields whose byte offsets fit 32 bits: site u has colour
// cb ^ (u & 1) (cb = (j + k) & 1 of the row) and packed index m + (u >> 1).
// Three 16-byte loads and one 2-byte load per colour; the colour is chosen by
// address, so every register index is fixed at compile time.
struct Row4 {
    double x[4], y[4], z[4];
    int s[4];
};

__device__ inline void load_row4(const Fields &F, uint32_t nsub, uint32_t m, uint32_t cb, Row4 &R)
{
    const uint32_t qa = cb * nsub + m, qb = (cb ^ 1u) * nsub + m;
    const double2 ax = fld2<true>(F.nx, qa), ay = fld2<true>(F.ny, qa), az = fld2<true>(F.nz, qa);
    const double2 bx = fld2<true>(F.nx, qb), by = fld2<true>(F.ny, qb), bz = fld2<true>(F.nz, qb);
    const uint32_t sa = spin2<true>(F.s, qa), sb = spin2<true>(F.s, qb);
    R.x[0] = ax.x; R.x[1] = bx.x; R.x[2] = ax.y; R.x[3] = bx.y;
    R.y[0] = ay.x; R.y[1] = by.x; R.y[2] = ay.y; R.y[3] = by.y;
    R.z[0] = az.x; R.z[1] = bz.x; R.z[2] = az.y; R.z[3] = bz.y;
    R.s[0] = lo_spin(sa); R.s[1] = lo_spin(sb); R.s[2] = hi_spin(sa); R.s[3] = hi_spin(sb);
}

// The observables of the four sites of packed indices m0, m0 + 1 (m0 even;
// with half even both lie in one row (j, k), packed offset h0 in the row)
// from its own row O, the +y row Y and the +z plane's row Z (Row4 each,
// Z unused without a +z bond).  The +x neighbour of site 3 and the nx of its
// diagonal site (i + 1, j + 1) are site 0 of the next pair's own and +y rows:
// taken from lane + 1 by DPP when nxt (that lane holds the next pair of this
// row in this pass), else read; past the row end i wraps to 0.  Same bond
// terms as measure_site, summed per site in x order.
template <int BC>
__device__ inline void measure_rows(const Geom &G, const Fields &F, double g, double KK, uint32_t m0, uint32_t h0,
                                    uint32_t j, uint32_t k, bool nxt, const Row4 &O, const Row4 &Y, const Row4 &Z,
                                    MeasureAcc &A)
{
    const uint32_t jp = (j + 1 == G.nj) ? 0u : j + 1;
    const bool zin = k + 1 < G.nk;
    const bool hz = zin || BC == BC_PERIODIC;
    const uint32_t my = m0 - j * G.half + jp * G.half;
    const uint32_t cb = (j + k) & 1u, cy = (jp + k) & 1u;
    double ex = lane_next(O.x[0]), ey = lane_next(O.y[0]), ez = lane_next(O.z[0]);
    int es = (int)lane_next((uint32_t)O.s[0]);
    double edx = lane_next(Y.x[0]);
    const bool wrap = h0 + 2u == G.half;
    if (!nxt || wrap) {
        const uint32_t mx = wrap ? m0 - h0 : m0 + 2u, mxy = wrap ? my - h0 : my + 2u;
        const uint32_t q = cb * G.nsub + mx;
        ex = F.nx[q]; ey = F.ny[q]; ez = F.nz[q]; es = F.s[q];
        edx = F.nx[cy * G.nsub + mxy];
    }
    const bool hy = (BC == BC_PERIODIC) || (j + 1 < G.nj);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const double ax = O.x[u], ay = O.y[u], az = O.z[u];
        const int sa = O.s[u];
        const double bx0 = u < 3 ? O.x[u + 1 < 4 ? u + 1 : 3] : ex;
        const double by0 = u < 3 ? O.y[u + 1 < 4 ? u + 1 : 3] : ey;
        const double bz0 = u < 3 ? O.z[u + 1 < 4 ? u + 1 : 3] : ez;
        const int sx = u < 3 ? O.s[u + 1 < 4 ? u + 1 : 3] : es;
        const double dgx = u < 3 ? Y.x[u + 1 < 4 ? u + 1 : 3] : edx;
        const double bx1 = Y.x[u], by1 = Y.y[u], bz1 = Y.z[u];
        const int sy = Y.s[u];
        const double tx = bond_term(ax, ay, az, bx0, by0, bz0, 0, sa, sx, KK);
        // the reference's +y bond: its dot product takes nx of the diagonal site
        double ety;
        {
            const double dd = ax * dgx + ay * by1 + az * bz1;
            const double cr = az * bx1 - ax * bz1;
            const double sf = 0.5 * (double)(sa + sy);
            ety = (1.0 - dd * dd) - KK * dd * cr * sf;
        }
        const double ty = bond_term(ax, ay, az, bx1, by1, bz1, 1, sa, sy, KK);
        double tz = 0.0, tzp = 0.0;
        if (hz) {
            tzp = bond_term(ax, ay, az, Z.x[u], Z.y[u], Z.z[u], 2, sa, Z.s[u], KK);
            if (zin) tz = tzp;
        }
        const double gs = g * (double)sa;
        A.v[0] += ((tx + ety) + tz) - gs;
        const bool hx = (BC == BC_PERIODIC) || u < 3 || !wrap;
        A.v[1] += ((hx ? tx : 0.0) + (hy ? ty : 0.0) + tzp) - gs;
        A.v[2] += ax * ax;
        A.v[3] += ay * ay;
        A.v[4] += az * az;
        A.v[5] += ax * ay;
        A.v[6] += ax * az;
        A.v[7] += ay * az;
        A.s += sa;
    }
}

// measure_rows on rows loaded here (the ensemble's form: one pass per row).
template <int BC>
__device__ inline void measure_pair(const Geom &G, const Fields &F, double g, double KK, uint32_t m0, bool nxt,
                                    MeasureAcc &A)
{
    const uint32_t row = fdiv(m0, G.fd_half);
    const uint32_t h0 = m0 - row * G.half;
    const uint32_t k = fdiv(row, G.fd_nj);
    const uint32_t j = row - k * G.nj;
    const uint32_t jp = (j + 1 == G.nj) ? 0u : j + 1;
    const bool hz = k + 1 < G.nk || BC == BC_PERIODIC;
    const uint32_t kp = k + 1 < G.nk ? k + 1 : 0u;
    Row4 O, Y, Z;
    load_row4(F, G.nsub, m0, (j + k) & 1u, O);
    load_row4(F, G.nsub, m0 - j * G.half + jp * G.half, (jp + k) & 1u, Y);
    if (hz) load_row4(F, G.nsub, m0 - k * G.plane + kp * G.plane, (j + kp) & 1u, Z);
    measure_rows<BC>(G, F, g, KK, m0, h0, j, k, nxt, O, Y, Z, A);
}

// The observables of one column of four-site rows through the planes
// [k0, k1) of a slab, marching along k: pair qd of a plane (packed indices
// 2 qd, 2 qd + 1 of each colour), up when `up`, else down.  Every step reads
// the +y row and one new plane row and carries the other from the step before
// -- marching up, the +z row becomes the next step's own row; marching down,
// the own row becomes the next step's +z row -- so each plane row is read
// once per slab.  Marching up, the +y row of the next step is read with this
// step's +z row, when the lanes of row j + 1 read that same row as their +z
// row.  nxt: the lane after this one holds the next pair of the plane.
template <int BC>
__device__ inline void measure_march_column(const Geom &G, const Fields &F, double g, double KK, uint32_t qd,
                                            uint32_t k0, uint32_t k1, bool up, bool nxt, MeasureAcc &acc)
{
    const uint32_t mp = 2u * qd;  // packed index of the pair in plane 0
    const uint32_t j = fdiv(mp, G.fd_half);
    const uint32_t h0 = mp - j * G.half;
    const uint32_t jp = (j + 1 == G.nj) ? 0u : j + 1;
    const uint32_t yo = jp * G.half - j * G.half;  // +y row offset (mod 2^32)
    Row4 O, Y, Z;
    if (up) {
        load_row4(F, G.nsub, k0 * G.plane + mp, (j + k0) & 1u, O);
        load_row4(F, G.nsub, k0 * G.plane + mp + yo, (jp + k0) & 1u, Y);
        for (uint32_t k = k0; k < k1; ++k) {
            const uint32_t m0 = k * G.plane + mp;
            const bool hz = k + 1 < G.nk || BC == BC_PERIODIC;
            const uint32_t kp = k + 1 < G.nk ? k + 1 : 0u;
            Row4 Yn;
            if (hz) load_row4(F, G.nsub, kp * G.plane + mp, (j + kp) & 1u, Z);
            if (k + 1 < k1) load_row4(F, G.nsub, m0 + G.plane + yo, (jp + k + 1) & 1u, Yn);
            measure_rows<BC>(G, F, g, KK, m0, h0, j, k, nxt, O, Y, Z, acc);
            O = Z;
            Y = Yn;
        }
    } else {
        {
            const uint32_t kp = k1 < G.nk ? k1 : 0u;
            if (k1 < G.nk || BC == BC_PERIODIC) load_row4(F, G.nsub, kp * G.plane + mp, (j + kp) & 1u, Z);
        }
        for (uint32_t k = k1; k-- > k0;) {
            const uint32_t m0 = k * G.plane + mp;
            load_row4(F, G.nsub, m0, (j + k) & 1u, O);
            load_row4(F, G.nsub, m0 + yo, (jp + k) & 1u, Y);
            measure_rows<BC>(G, F, g, KK, m0, h0, j, k, nxt, O, Y, Z, acc);
            Z = O;
        }
    }
}

// Deterministic block reduction of MeasureAcc: wave butterflies, then a
// fixed-order pass over the waves; writes 8 doubles to part and one int to spart.
template <int BS>
__device__ inline void block_reduce_measure(MeasureAcc &A, double *part, long long *spart)
{
    __shared__ double wv[BS / 64][8];
    __shared__ long long ws[BS / 64];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int q = 0; q < 8; ++q) A.v[q] += __shfl_xor(A.v[q], off, 64);
        A.s += __shfl_xor(A.s, off, 64);
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) wv[wid][q] = A.v[q];
        ws[wid] = A.s;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        double t = 0.0;
        for (int w = 0; w < BS / 64; ++w) t += wv[w][threadIdx.x];
        part[threadIdx.x] = t;
    }
    if (threadIdx.x == 8) {
        long long t = 0;
        for (int w = 0; w < BS / 64; ++w) t += ws[w];
        *spart = t;
    }
}

}  // namespace chll
