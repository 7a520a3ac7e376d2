/*
 * chll_oracle.c -- CPU restatement of the chiral Lebwohl-Lasher Metropolis
 * sweep of mdeutsc2/ChLLModel (chll/chll_mod.f90), written in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library
 * (chllmodel_amd/csrc) is compared against; nothing in the product links or
 * loads it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may use it.
 *
 * Arithmetic contract ("compat" arithmetic, the reference compiled without FMA
 * contraction): every operation is binary64, evaluated left to right exactly as
 * the Fortran source writes it, except that the site energy is a default REAL
 * (binary32) function, so its running value is rounded to binary32 after every
 * bond and after the field term (chll/chll_mod.f90:363, 369-411).  Build with
 * -ffp-contract=off so no fused multiply-add changes a rounding.
 *
 * The "fp64" arithmetic variant drops the binary32 roundings (energies stay
 * binary64); it is the oracle for the library's pure-FP64 mode.
 *
 * Uniform sources.  The reference draws host uniforms with the Fortran
 * intrinsic random_number (chll/chll_mod.f90:58, 66-80), whose stream is
 * compiler-defined and unseeded.  The oracle replaces it by one of:
 *   ORACLE_RNG_STREAM : a documented counter-based stream u_n (Philox4x32-10,
 *                       below) consumed in the reference's exact draw schedule;
 *   ORACLE_RNG_BUFFER : an explicit caller-supplied array consumed in the same
 *                       schedule;
 *   ORACLE_RNG_PHILOX : the library's production, site-keyed Philox counters
 *                       (one counter block per site, sweep and colour; see
 *                       philox_site), so the production GPU path can be
 *                       compared with the oracle draw for draw.
 *
 * Indices: 0-based (i, j, k), Fortran (column-major) linear order
 * l = i + ni*(j + nj*k), as the reference's arrays nx(ni,nj,nk) etc.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_RNG_STREAM 0
#define ORACLE_RNG_BUFFER 1
#define ORACLE_RNG_PHILOX 2

#define ORACLE_BC_OPEN 0
#define ORACLE_BC_PERIODIC 1

#define ORACLE_ARITH_COMPAT 0
#define ORACLE_ARITH_FP64 1

/* 2*pi as the reference forms it: 2 * (4*atan(1d0)) (chll/chll_mod.f90:127-128). */
static const double TWOPI = 6.283185307179586;
/* The single-precision literal 0.995 widened to binary64 (chll/chll_mod.f90:86,88). */
static const double STEP_SHRINK = 0.99500000476837158203125;

static inline double fl32(double x) { return (double)(float)x; }

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11), restated from the published algorithm.  Pinned by the
 * published known-answer vectors in tests/. */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0,1) from two 32-bit words (hi word first). */
static inline double u53(uint32_t hi, uint32_t lo)
{
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return (double)(x >> 11) * 0x1.0p-53;
}

/* Deterministic-mode stream: u_n, n = 0, 1, 2, ...  Block q = n/2 is
 * Philox4x32-10 with counter (q_lo, q_hi, 0, 0x5354524D) and key (seed_lo,
 * seed_hi); even n takes words (1,0), odd n takes words (3,2). */
double oracle_stream_uniform(uint64_t seed, uint64_t n)
{
    uint64_t q = n >> 1;
    uint32_t ctr[4] = {(uint32_t)q, (uint32_t)(q >> 32), 0u, 0x5354524Du};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    oracle_philox4x32_10(ctr, key, w);
    return (n & 1) ? u53(w[3], w[2]) : u53(w[1], w[0]);
}

void oracle_stream_fill(uint64_t seed, uint64_t n0, int64_t count, double *out)
{
    for (int64_t t = 0; t < count; ++t) out[t] = oracle_stream_uniform(seed, n0 + (uint64_t)t);
}

/* Production site-keyed draws: one block per site, sweep and colour.  Counter
 * (m, istep, (run_index << 4) | (colour << 1), replica), key (seed_lo, seed_hi);
 * m = packed index of the site within its colour (l >> 1).  Director trial of
 * sweep t: block t, u_phi = word 0 / 2^32, u_acc = word 1 and the high half of
 * word 2 as a 48-bit fraction.  Chirality trial of sweep t: word 3 and the low
 * half of word 2 as a 48-bit fraction, of block t for colour 1 and of block
 * t + 1 for colour 0. */
static inline double u48(uint32_t hi, uint32_t lo16)
{
    return (double)hi / 4294967296.0 + (double)lo16 / 281474976710656.0;
}

static void philox_site(uint64_t seed, uint32_t run_index, uint32_t replica, uint32_t m,
                        uint32_t istep, int colour, double *u_phi, double *u_dir, double *u_spin)
{
    uint32_t ctr[4] = {m, istep, (run_index << 4) | ((uint32_t)colour << 1), replica};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    oracle_philox4x32_10(ctr, key, w);
    *u_phi = (double)w[0] / 4294967296.0;
    *u_dir = u48(w[1], w[2] >> 16);
    *u_spin = u48(w[3], w[2] & 0xffffu);
}

/* Initial states keyed by (seed, replica, site) (the library's device
 * generators; the reference's own are Python loops over numpy's global
 * stream: scripts/pdanneal.py:5-16, scripts/hero480.py:18-21, chll/chll.py:
 * 84-98).  Site of Fortran index l: Philox blocks with counters
 * (l, replica, 0x494E4954, 0) and (..., 1), key (seed_lo, seed_hi).
 * kind 0: costh = 2 (u53(w0, w1) - 0.5), phi = 2 pi u53(w2, w3),
 * n = (sqrt(1 - costh^2) cos phi, sqrt(1 - costh^2) sin phi, costh);
 * kind 1: n = (1, 0, 0); kind 2: n = (1, 0, 1) / sqrt(2).
 * Spin: -1 when u53(v0, v1) < p_flip, else +1.  Fortran-ordered arrays. */
void oracle_init_state(double *nx, double *ny, double *nz, int32_t *s, int ni, int nj, int nk, int kind,
                       uint64_t seed, uint32_t replica, double p_flip)
{
    const int64_t n3 = (int64_t)ni * nj * nk;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int64_t l = 0; l < n3; ++l) {
        uint32_t c0[4] = {(uint32_t)l, replica, 0x494E4954u, 0u}, c1[4] = {(uint32_t)l, replica, 0x494E4954u, 1u};
        uint32_t w[4], v[4];
        if (kind == 0) {
            oracle_philox4x32_10(c0, key, w);
            double costh = 2.0 * (u53(w[0], w[1]) - 0.5);
            double sinth = sqrt(1.0 - costh * costh);
            double phi = u53(w[2], w[3]) * TWOPI;
            nx[l] = sinth * cos(phi);
            ny[l] = sinth * sin(phi);
            nz[l] = costh;
        } else if (kind == 1) {
            nx[l] = 1.0; ny[l] = 0.0; nz[l] = 0.0;
        } else {
            nx[l] = 0.70710678118654752440; ny[l] = 0.0; nz[l] = 0.70710678118654752440;
        }
        oracle_philox4x32_10(c1, key, v);
        s[l] = u53(v[0], v[1]) < p_flip ? -1 : 1;
    }
}

/* ------------------------------------------------------------------------- */
/* Sublattice lists (chll/chll_mod.f90:9-38): Fortran order, 1-based (i+j+k)
 * odd -> sl1, even -> sl2; entries are 1-based linear indices. */
void oracle_init(int32_t *sl1, int32_t *sl2, int ni, int nj, int nk)
{
    int64_t n1 = 0, n2 = 0;
    for (int k = 1; k <= nk; ++k)
        for (int j = 1; j <= nj; ++j)
            for (int i = 1; i <= ni; ++i) {
                int32_t idx = i + (j - 1) * ni + (k - 1) * (ni * nj);
                if ((i + j + k) % 2 != 0) sl1[n1++] = idx;
                else sl2[n2++] = idx;
            }
}

/* ------------------------------------------------------------------------- */
typedef struct {
    double *nx, *ny, *nz;
    int32_t *s;
    int ni, nj, nk;
    int bc;
    int arith;
} lat_t;

static inline int64_t lidx(const lat_t *L, int i, int j, int k)
{
    return (int64_t)i + (int64_t)L->ni * ((int64_t)j + (int64_t)L->nj * k);
}

/* Neighbour index along one axis, or -1 when absent (open boundary,
 * chll/chll_mod.f90:136-141, 371-410). */
static inline int nb(int x, int dx, int n, int bc)
{
    int y = x + dx;
    if (y >= 0 && y < n) return y;
    if (bc == ORACLE_BC_PERIODIC) return (y + n) % n;
    return -1;
}

/* One bond update of the site energy; dir = +1 for the +x/+y/+z bonds
 * (the chiral term is subtracted) and -1 for the -x/-y/-z bonds (added)
 * (chll/chll_mod.f90:371-410).  axis selects the cross-product component. */
static inline double bond_update(double e, int arith, int axis, int dir,
                                 double nnx, double nny, double nnz,
                                 double mx, double my, double mz,
                                 int snew, int sb, double KK)
{
    double dott = nnx * mx + nny * my + nnz * mz;
    double cr;
    if (axis == 0) cr = nny * mz - nnz * my;
    else if (axis == 1) cr = nnz * mx - nnx * mz;
    else cr = nnx * my - nny * mx;
    double sfac = 0.5 * (double)(snew + sb);
    double r;
    if (dir > 0) r = e + (1.0 - dott * dott) - KK * dott * cr * sfac;
    else r = e + (1.0 - dott * dott) + KK * dott * cr * sfac;
    return arith == ORACLE_ARITH_COMPAT ? fl32(r) : r;
}

/* Site energy E(n, snew) at (i,j,k), neighbours read from the lattice
 * (chll/chll_mod.f90:363-412). */
static double site_energy(const lat_t *L, int i, int j, int k,
                          double nnx, double nny, double nnz, int snew, double g, double KK)
{
    double e = 0.0;
    const int bc = L->bc, ar = L->arith;
    const int di[6] = {+1, -1, 0, 0, 0, 0};
    const int dj[6] = {0, 0, +1, -1, 0, 0};
    const int dk[6] = {0, 0, 0, 0, +1, -1};
    for (int b = 0; b < 6; ++b) {
        int ii = i, jj = j, kk = k;
        if (di[b]) { ii = nb(i, di[b], L->ni, bc); if (ii < 0) continue; }
        if (dj[b]) { jj = nb(j, dj[b], L->nj, bc); if (jj < 0) continue; }
        if (dk[b]) { kk = nb(k, dk[b], L->nk, bc); if (kk < 0) continue; }
        int64_t q = lidx(L, ii, jj, kk);
        e = bond_update(e, ar, b >> 1, (b & 1) ? -1 : +1, nnx, nny, nnz,
                        L->nx[q], L->ny[q], L->nz[q], snew, L->s[q], KK);
    }
    double r = e - g * (double)snew;
    return ar == ORACLE_ARITH_COMPAT ? fl32(r) : r;
}

double oracle_site_energy(const double *nx, const double *ny, const double *nz, const int32_t *s,
                          int ni, int nj, int nk, int bc, int arith, int i, int j, int k,
                          double nnx, double nny, double nnz, int snew, double g, double KK)
{
    lat_t L = {(double *)nx, (double *)ny, (double *)nz, (int32_t *)s, ni, nj, nk, bc, arith};
    return site_energy(&L, i, j, k, nnx, nny, nnz, snew, g, KK);
}

/* Cone trial direction (chll/chll_mod.f90:149-183). */
static void trial_director(double nnx, double nny, double nnz, double d, double u_phi,
                           double *ox, double *oy, double *oz)
{
    double phi = u_phi * TWOPI;
    double cphi, sphi;
    sincos(phi, &sphi, &cphi);
    if (fabs(nnz) > 0.999) {
        double xx = nnx + d * cphi;
        double yy = nny + d * sphi;
        double zz = nnz;
        double rsq = sqrt(xx * xx + yy * yy + zz * zz);
        *ox = xx / rsq; *oy = yy / rsq; *oz = zz / rsq;
    } else {
        double ux = -nny, uy = nnx, uz = 0.0;
        double rsq = sqrt(ux * ux + uy * uy);
        ux = ux / rsq; uy = uy / rsq; uz = uz / rsq;
        double vx = -nnz * nnx, vy = -nnz * nny, vz = nnx * nnx + nny * nny;
        rsq = sqrt(vx * vx + vy * vy + vz * vz);
        vx = vx / rsq; vy = vy / rsq; vz = vz / rsq;
        double dc = d * cphi, ds = d * sphi;
        double tx = nnx + dc * ux + ds * vx;
        double ty = nny + dc * uy + ds * vy;
        double tz = nnz + dc * uz + ds * vz;
        rsq = sqrt(tx * tx + ty * ty + tz * tz);
        *ox = tx / rsq; *oy = ty / rsq; *oz = tz / rsq;
    }
}

void oracle_trial_director(double nnx, double nny, double nnz, double d, double u_phi, double *out3)
{
    trial_director(nnx, nny, nnz, d, u_phi, &out3[0], &out3[1], &out3[2]);
}

/* Metropolis rule (chll/chll_mod.f90:188-200, 235-243). */
static inline int metropolis(double enew, double eold, double u, double kbt)
{
    if (enew < eold) return 1;
    return u <= exp(-(enew - eold) / kbt);
}

/* ------------------------------------------------------------------------- */
/* Uniform source in the reference's per-sweep draw schedule
 * (chll/chll_mod.f90:66-80): chunks of nsub, 2nsub, nsub, 2nsub, 2nsub, 2nsub
 * uniforms per sweep; the 2nsub chunks are rand2(nsub,2) in column-major
 * order. */
typedef struct {
    int mode;
    uint64_t seed;
    uint64_t pos;          /* next draw index (stream / buffer) */
    const double *buf;
    int64_t buf_len;
    uint32_t run_index, replica;
    uint32_t site_offset;  /* added to the packed index m of the Philox counter */
    int overflow;
} usrc_t;

static inline double src_at(usrc_t *U, uint64_t n)
{
    if (U->mode == ORACLE_RNG_STREAM) return oracle_stream_uniform(U->seed, n);
    if ((int64_t)n >= U->buf_len) { __atomic_store_n(&U->overflow, 1, __ATOMIC_RELAXED); return 2.0; }
    return U->buf[n];
}

/* ------------------------------------------------------------------------- */
/* Rank of site (i, j, k) in its sublattice list sl1 (colour 0) or sl2 (colour
 * 1), i.e. the index r of its uniforms rand1(r+1), rand2(r+1, :) in the
 * reference's draw schedule (chll/chll_mod.f90:23-36, 131-134, 146, 188).
 * The lists hold the sites of one colour in Fortran order, so r counts the
 * sites of the same colour before linear index l = i + ni (j + nj k).  For
 * every shape with N even, r = l >> 1: stepping l -> l + 1 changes i + j + k
 * by 1 inside a row, by 2 - ni at a row end and by 3 - ni - nj at a plane end,
 * so the colour alternates along l inside rows, across row ends when ni is
 * odd and across plane ends when ni + nj is even.  Where it does not, the
 * row (ni even) or plane (ni odd, nj even) just ended holds an even number of
 * sites, half of each colour, and the next one starts at an even l.  Either
 * way the sites of one colour before l number l >> 1.  (Checked exhaustively
 * against the lists of oracle_init in the tests.) */
int64_t oracle_rank(int ni, int nj, int i, int j, int k)
{
    return ((int64_t)i + (int64_t)ni * ((int64_t)j + (int64_t)nj * k)) >> 1;
}

/* ------------------------------------------------------------------------- */
/* Director half-sweep over one colour (chll/chll_mod.f90:117-202).  colour 0 is
 * sl1 (1-based i+j+k odd = 0-based even), colour 1 is sl2.  The sublattice rank
 * of the site (oracle_rank) indexes rand1 / rand2 and the Philox counter. */
static int64_t evolve_dir(lat_t *L, int colour, usrc_t *U, uint64_t base_phi, uint64_t base_acc,
                          uint32_t istep, double g, double KK, double d, double kbt, int nthreads)
{
    int64_t nacc = 0;
    /* Sites of one colour share no bond, so the planes may run in any order
     * and on any number of threads with the same result. */
#pragma omp parallel for schedule(static) reduction(+ : nacc) num_threads(nthreads) if (nthreads > 1)
    for (int k = 0; k < L->nk; ++k)
        for (int j = 0; j < L->nj; ++j)
            for (int i = 0; i < L->ni; ++i) {
                if (((i + j + k) & 1) != colour) continue;
                int64_t l = lidx(L, i, j, k);
                int64_t r = oracle_rank(L->ni, L->nj, i, j, k);
                double u_phi, u_acc, unused;
                if (U->mode == ORACLE_RNG_PHILOX) {
                    philox_site(U->seed, U->run_index, U->replica, (uint32_t)r + U->site_offset, istep, colour,
                                &u_phi, &u_acc, &unused);
                } else {
                    u_phi = src_at(U, base_phi + (uint64_t)r);
                    u_acc = src_at(U, base_acc + (uint64_t)r);
                }
                double nnx = L->nx[l], nny = L->ny[l], nnz = L->nz[l];
                int sl = L->s[l];
                double eold = site_energy(L, i, j, k, nnx, nny, nnz, sl, g, KK);
                double tx, ty, tz;
                trial_director(nnx, nny, nnz, d, u_phi, &tx, &ty, &tz);
                double enew = site_energy(L, i, j, k, tx, ty, tz, sl, g, KK);
                if (metropolis(enew, eold, u_acc, kbt)) {
                    L->nx[l] = tx; L->ny[l] = ty; L->nz[l] = tz;
                    ++nacc;
                }
            }
    return nacc;
}

/* Chirality half-sweep over one colour (chll/chll_mod.f90:204-245). */
static int64_t evolve_s(lat_t *L, int colour, usrc_t *U, uint64_t base_acc,
                        uint32_t istep, double g, double KK, double kbt, int nthreads)
{
    int64_t nflip = 0;
#pragma omp parallel for schedule(static) reduction(+ : nflip) num_threads(nthreads) if (nthreads > 1)
    for (int k = 0; k < L->nk; ++k)
        for (int j = 0; j < L->nj; ++j)
            for (int i = 0; i < L->ni; ++i) {
                if (((i + j + k) & 1) != colour) continue;
                int64_t l = lidx(L, i, j, k);
                int64_t r = oracle_rank(L->ni, L->nj, i, j, k);
                double u_acc, unused0, unused1;
                if (U->mode == ORACLE_RNG_PHILOX)
                    philox_site(U->seed, U->run_index, U->replica, (uint32_t)r + U->site_offset,
                                colour ? istep : istep + 1u, colour, &unused0, &unused1, &u_acc);
                else
                    u_acc = src_at(U, base_acc + (uint64_t)r);
                double nnx = L->nx[l], nny = L->ny[l], nnz = L->nz[l];
                int sl = L->s[l];
                double eold = site_energy(L, i, j, k, nnx, nny, nnz, sl, g, KK);
                int snew = -sl;
                double enew = site_energy(L, i, j, k, nnx, nny, nnz, snew, g, KK);
                if (metropolis(enew, eold, u_acc, kbt)) {
                    L->s[l] = snew;
                    ++nflip;
                }
            }
    return nflip;
}

/* ------------------------------------------------------------------------- */
/* Reported total energy per site (chll/chll_mod.f90:414-456): periodic in i and
 * j, open in k, and the +y bond's x-component taken from the diagonal site
 * (ip1, jp1, k) as the reference writes it (:438).  binary64 serial sum in
 * (k, j, i) order, divided by REAL(4) N, rounded to binary32. */
double oracle_etot(const double *nx, const double *ny, const double *nz, const int32_t *s,
                   int ni, int nj, int nk, double g, double KK)
{
    double e = 0.0;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                int ip = (i + 1) % ni, jp = (j + 1) % nj;
                int64_t c = (int64_t)i + (int64_t)ni * ((int64_t)j + (int64_t)nj * k);
                int64_t px = (int64_t)ip + (int64_t)ni * ((int64_t)j + (int64_t)nj * k);
                int64_t py = (int64_t)i + (int64_t)ni * ((int64_t)jp + (int64_t)nj * k);
                int64_t pd = (int64_t)ip + (int64_t)ni * ((int64_t)jp + (int64_t)nj * k);
                double ddot = nx[c] * nx[px] + ny[c] * ny[px] + nz[c] * nz[px];
                double cr = ny[c] * nz[px] - nz[c] * ny[px];
                double sfac = 0.5 * (double)(s[c] + s[px]);
                e = e + (1.0 - ddot * ddot) - KK * ddot * cr * sfac;

                ddot = nx[c] * nx[pd] + ny[c] * ny[py] + nz[c] * nz[py];
                cr = nz[c] * nx[py] - nx[c] * nz[py];
                sfac = 0.5 * (double)(s[c] + s[py]);
                e = e + (1.0 - ddot * ddot) - KK * ddot * cr * sfac;

                if (k + 1 < nk) {
                    int64_t pz = c + (int64_t)ni * nj;
                    ddot = nx[c] * nx[pz] + ny[c] * ny[pz] + nz[c] * nz[pz];
                    cr = nx[c] * ny[pz] - ny[c] * nx[pz];
                    sfac = 0.5 * (double)(s[c] + s[pz]);
                    e = e + (1.0 - ddot * ddot) - KK * ddot * cr * sfac;
                }
                e = e - g * (double)s[c];
            }
    int64_t n3 = (int64_t)ni * nj * nk;
    return fl32(e / (double)(float)n3);
}

/* Lattice Hamiltonian that the sweep samples, each bond counted once, in plain
 * binary64 (the quantity the library reports as its "true" energy): sum over
 * sites of the +x, +y, +z bonds that exist under bc, minus g * sum(s). */
double oracle_energy_true(const double *nx, const double *ny, const double *nz, const int32_t *s,
                          int ni, int nj, int nk, int bc, double g, double KK)
{
    double e = 0.0;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                int64_t c = (int64_t)i + (int64_t)ni * ((int64_t)j + (int64_t)nj * k);
                int nbs[3][3] = {{nb(i, 1, ni, bc), j, k}, {i, nb(j, 1, nj, bc), k}, {i, j, nb(k, 1, nk, bc)}};
                for (int a = 0; a < 3; ++a) {
                    if (nbs[a][0] < 0 || nbs[a][1] < 0 || nbs[a][2] < 0) continue;
                    int64_t q = (int64_t)nbs[a][0] + (int64_t)ni * ((int64_t)nbs[a][1] + (int64_t)nj * nbs[a][2]);
                    double dott = nx[c] * nx[q] + ny[c] * ny[q] + nz[c] * nz[q];
                    double cr;
                    if (a == 0) cr = ny[c] * nz[q] - nz[c] * ny[q];
                    else if (a == 1) cr = nz[c] * nx[q] - nx[c] * nz[q];
                    else cr = nx[c] * ny[q] - ny[c] * nx[q];
                    double sfac = 0.5 * (double)(s[c] + s[q]);
                    e += (1.0 - dott * dott) - KK * dott * cr * sfac;
                }
                e -= g * (double)s[c];
            }
    return e;
}

/* ------------------------------------------------------------------------- */
/* The sweep driver (chll/chll_mod.f90:40-115) with the uniform source above.
 * Phase order per sweep: director(sl1), director(sl2), chirality(sl1),
 * chirality(sl2) (:70-81).  Step adaptation (:83-90) and measurement rows
 * (:91-105) as the reference computes them; m is column-major [rows,5].
 * Rows beyond rows = (1+nstop-nstart)/nout are not written.  site_offset is
 * added to every Philox site counter: the lattice is then the slab of planes
 * k0.. of a larger lattice with the same ni, nj, whose packed indices start at
 * site_offset = ni nj k0 / 2 (k0 even keeps the colours).
 * Returns 0, or -1 when a buffer source ran out. */
static int run_impl(int rng_mode, uint64_t seed, uint64_t *stream_pos, const double *buf, int64_t buf_len,
                    uint32_t run_index, uint32_t replica, uint32_t site_offset, int bc, int arith,
                    int nstart, int nstop, int nout,
                    double *nx, double *ny, double *nz, int32_t *s,
                    double g, double KK, double *d, double kbt,
                    int ni, int nj, int nk, int32_t *msteps, double *m, int64_t *counts,
                    int nthreads, int hostfill)
{
    lat_t L = {nx, ny, nz, s, ni, nj, nk, bc, arith};
    usrc_t U;
    memset(&U, 0, sizeof(U));
    U.mode = rng_mode; U.seed = seed; U.pos = stream_pos ? *stream_pos : 0;
    U.buf = buf; U.buf_len = buf_len; U.run_index = run_index; U.replica = replica;
    U.site_offset = site_offset;
    int64_t n3 = (int64_t)ni * nj * nk;
    int64_t nsub = n3 / 2;
    int rows = (1 + nstop - nstart) / nout;
    int mc = 0;
    double dd = *d;
    /* hostfill (stream draws): every sweep first fills all 10 nsub uniforms of
     * its six random_number calls into a host array, as the reference does
     * (chll/chll_mod.f90:66-80), and the half-sweeps read them from there */
    double *fb = NULL;
    usrc_t F = U;
    if (hostfill && rng_mode == ORACLE_RNG_STREAM) {
        fb = (double *)malloc((size_t)(10 * nsub > 0 ? 10 * nsub : 1) * sizeof(double));
        if (!fb) return -2;
        F.mode = ORACLE_RNG_BUFFER; F.buf = fb; F.buf_len = 10 * nsub;
    }
    for (int istep = nstart; istep <= nstop; ++istep) {
        uint64_t b = U.pos;
        usrc_t *src = &U;
        if (fb) {
            const int64_t nd = 10 * nsub;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
            for (int64_t t = 0; t < nd; ++t) fb[t] = oracle_stream_uniform(seed, b + (uint64_t)t);
            src = &F;
            b = 0;
        }
        /* chunk offsets of the six random_number calls of one sweep */
        uint64_t c0 = b, c1 = b + nsub, c2 = b + 3 * nsub, c3 = b + 4 * nsub, c4 = b + 6 * nsub, c5 = b + 8 * nsub;
        int64_t nacc = 0, nflip = 0;
        nacc += evolve_dir(&L, 0, src, c0, c1, (uint32_t)istep, g, KK, dd, kbt, nthreads);
        nacc += evolve_dir(&L, 1, src, c2, c3, (uint32_t)istep, g, KK, dd, kbt, nthreads);
        nflip += evolve_s(&L, 0, src, c4 + nsub, (uint32_t)istep, g, KK, kbt, nthreads);
        nflip += evolve_s(&L, 1, src, c5 + nsub, (uint32_t)istep, g, KK, kbt, nthreads);
        b = U.pos;
        if (U.mode != ORACLE_RNG_PHILOX) U.pos = b + 10 * (uint64_t)nsub;
        if (counts) { counts[2 * (istep - nstart)] = nacc; counts[2 * (istep - nstart) + 1] = nflip; }
        double paccept = (double)((float)(int32_t)nacc / (float)n3);
        if (istep < 10000) {
            if (paccept < 0.4) dd = dd * STEP_SHRINK;
            else if (paccept > 0.6) dd = dd / STEP_SHRINK;
        }
        if (istep % nout == 0) {
            double faccept = (double)((float)(int32_t)nflip / (float)n3);
            double te = oracle_etot(nx, ny, nz, s, ni, nj, nk, g, KK);
            float ssum = 0.0f;
            for (int64_t q = 0; q < n3; ++q) ssum += (float)s[q];
            double ex = (double)(ssum / (float)n3);
            if (mc < rows) {
                msteps[mc] = istep;
                m[mc + 0 * rows] = te;
                m[mc + 1 * rows] = ex;
                m[mc + 2 * rows] = paccept;
                m[mc + 3 * rows] = faccept;
                m[mc + 4 * rows] = dd;
            }
            ++mc;
        }
    }
    free(fb);
    *d = dd;
    if (stream_pos) *stream_pos = U.pos;
    return U.overflow ? -1 : 0;
}

int oracle_run(int rng_mode, uint64_t seed, uint64_t *stream_pos, const double *buf, int64_t buf_len,
               uint32_t run_index, uint32_t replica, uint32_t site_offset, int bc, int arith,
               int nstart, int nstop, int nout,
               double *nx, double *ny, double *nz, int32_t *s,
               double g, double KK, double *d, double kbt,
               int ni, int nj, int nk, int32_t *msteps, double *m,
               int64_t *counts /* optional [nsweeps*2]: naccept, nflip per sweep */)
{
    return run_impl(rng_mode, seed, stream_pos, buf, buf_len, run_index, replica, site_offset, bc, arith, nstart,
                    nstop, nout, nx, ny, nz, s, g, KK, d, kbt, ni, nj, nk, msteps, m, counts, 1, 0);
}

/* oracle_run on nthreads threads (OpenMP over the planes of each half-sweep,
 * the analogue of the reference's -acc=multicore build, chll/Makefile:8,10);
 * hostfill as in run_impl.  The lattice, rows and counts are those of
 * oracle_run: sites of one colour are independent. */
int oracle_run_mt(int rng_mode, uint64_t seed, uint64_t *stream_pos, const double *buf, int64_t buf_len,
                  uint32_t run_index, uint32_t replica, uint32_t site_offset, int bc, int arith,
                  int nstart, int nstop, int nout,
                  double *nx, double *ny, double *nz, int32_t *s,
                  double g, double KK, double *d, double kbt,
                  int ni, int nj, int nk, int32_t *msteps, double *m, int64_t *counts, int nthreads, int hostfill)
{
    return run_impl(rng_mode, seed, stream_pos, buf, buf_len, run_index, replica, site_offset, bc, arith, nstart,
                    nstop, nout, nx, ny, nz, s, g, KK, d, kbt, ni, nj, nk, msteps, m, counts,
                    nthreads < 1 ? 1 : nthreads, hostfill);
}
