// chll_device.h -- device-side building blocks of the chiral Lebwohl-Lasher
// sweep for gfx950: lattice geometry of the colour-split layout, counter-based
// Philox4x32-10, the site-energy bond terms and the cone trial move.
//
// Arithmetic: this library is compiled with -ffp-contract=off.  The "compat"
// functions evaluate every binary64 operation in the order chll/chll_mod.f90
// writes it and round the running site energy to binary32 after every bond
// (the reference's `real function energy`, chll/chll_mod.f90:363-412), so a
// deterministic run reproduces the reference's accept/reject decisions.  The
// "fp64" functions keep energies in binary64, use explicit fused multiply-adds
// and evaluate energy differences directly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chll {

enum : int { ARITH_COMPAT = 0, ARITH_FP64 = 1 };
enum : int { RNG_PHILOX = 0, RNG_STREAM = 1, RNG_BUFFER = 2 };
enum : int { BC_OPEN = 0, BC_PERIODIC = 1 };

// 2*pi as the reference forms it, 2 * (4 * atan(1d0)) (chll/chll_mod.f90:127-128).
constexpr double kTwoPi = 6.283185307179586;

// Unsigned 32-bit division by a run-time constant d >= 1 (Granlund &
// Montgomery, "Division by invariant integers using multiplication", PLDI'94,
// Fig. 4.1 with N = 32): l = ceil(log2 d), mul = floor(2^32 (2^l - d) / d) + 1,
// q = (mulhi(mul, n) + n) >> l, the sum taken in 64 bits.  Exact for every
// n < 2^32.
struct FastDiv {
    uint32_t d, mul, shift;
};

__host__ __device__ inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((uint64_t(1) << l) < d) ++l;
    f.shift = l;
    f.mul = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
    return f;
}

__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv &f)
{
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t t = __umulhi(n, f.mul);
#else
    uint32_t t = (uint32_t)(((uint64_t)n * f.mul) >> 32);
#endif
    return (uint32_t)(((uint64_t)t + n) >> f.shift);
}

// Colour-split structure-of-arrays layout.  Site (i,j,k) (0-based, Fortran
// order l = i + ni*(j + nj*k)) has colour c = (i+j+k) & 1 -- colour 0 is the
// reference's sl1 (1-based i+j+k odd), colour 1 is sl2 (chll/chll_mod.f90:23-36)
// -- and packed index m = l >> 1 within its colour.  For every shape with N
// even, m is also the site's rank in its sublattice list, i.e. the index of
// its uniforms in rand1 / rand2: the colour alternates along l inside rows,
// across row ends when ni is odd and across plane ends when ni + nj is even;
// where it does not, the row (ni even) or plane (ni odd, nj even) just ended
// holds an even number of sites, half of each colour.  With ni even a row
// holds ni / 2 sites of each colour at consecutive m; with ni odd it holds
// (ni +- 1) / 2, and the kernels decode m through l (decode, neighbours).
// Each field stores colour 0 at [0, nsub) and colour 1 at [nsub, 2 nsub).
//
// Index 2 nsub of every field is a padding "zero site": director (0,0,0),
// spin 0.  Absent neighbours (open boundaries) point there, so a bond to it
// has c = q = 0 and adds nothing to a binary64 energy difference.
constexpr uint32_t kFieldPad = 64;  // padding elements after the 2 nsub sites

struct Geom {
    uint32_t ni, nj, nk;
    uint32_t half;   // ni / 2: packed sites per row and colour (ni even)
    uint32_t plane;  // ni * nj / 2: packed sites per plane and colour (ni even)
    uint32_t nsub;   // N / 2
    uint32_t zero;   // 2 nsub: index of the zero site
    uint32_t nplane; // ni * nj: sites per plane
    FastDiv fd_half, fd_nj, fd_ni;
};

struct Site {
    uint32_t i, j, k;
};

// Site of Fortran linear index l.
__device__ inline Site decode_l(const Geom &G, uint32_t l)
{
    const uint32_t row = fdiv(l, G.fd_ni);
    const uint32_t k = fdiv(row, G.fd_nj);
    return Site{l - row * G.ni, row - k * G.nj, k};
}

__device__ inline Site decode(const Geom &G, uint32_t m, int colour)
{
    if (G.ni & 1u) {
        // the site of this colour among l = 2m, 2m + 1 (adjacent in Fortran
        // order, of different colours)
        Site s = decode_l(G, 2u * m);
        if (((s.i + s.j + s.k) & 1u) != (uint32_t)colour) {
            if (++s.i == G.ni) {
                s.i = 0;
                if (++s.j == G.nj) { s.j = 0; ++s.k; }
            }
        }
        return s;
    }
    uint32_t row = fdiv(m, G.fd_half);
    uint32_t h = m - row * G.half;
    uint32_t k = fdiv(row, G.fd_nj);
    uint32_t j = row - k * G.nj;
    uint32_t p = (j + k + (uint32_t)colour) & 1u;
    return Site{2u * h + p, j, k};
}

// Field indices of the six neighbours (all of the other colour, whose block
// starts at oth) in the order +x, -x, +y, -y, +z, -z, and a presence mask
// (bit b set = bond b exists).  Open boundaries drop absent bonds
// (chll/chll_mod.f90:136-141, 371-410) and point them at the zero site;
// periodic boundaries wrap them.
template <int BC>
__device__ inline uint32_t neighbours(const Geom &G, uint32_t m, const Site &s, uint32_t oth, uint32_t nb[6])
{
    if (BC == BC_OPEN && (G.ni & 1u)) {
        // odd ni (open boundaries only): packed index of linear index l' is l' >> 1
        const uint32_t l = s.i + G.ni * (s.j + G.nj * s.k);
        uint32_t mask = 0;
        nb[0] = s.i + 1 < G.ni ? (mask |= 1u, oth + ((l + 1u) >> 1)) : G.zero;
        nb[1] = s.i >= 1 ? (mask |= 2u, oth + ((l - 1u) >> 1)) : G.zero;
        nb[2] = s.j + 1 < G.nj ? (mask |= 4u, oth + ((l + G.ni) >> 1)) : G.zero;
        nb[3] = s.j >= 1 ? (mask |= 8u, oth + ((l - G.ni) >> 1)) : G.zero;
        nb[4] = s.k + 1 < G.nk ? (mask |= 16u, oth + ((l + G.nplane) >> 1)) : G.zero;
        nb[5] = s.k >= 1 ? (mask |= 32u, oth + ((l - G.nplane) >> 1)) : G.zero;
        return mask;
    }
    const uint32_t p = s.i & 1u;
    const uint32_t o = oth + m;
    uint32_t mask = 0;
    // +x
    if (s.i + 1 < G.ni) { nb[0] = o + p; mask |= 1u; }
    else if (BC == BC_PERIODIC) { nb[0] = o + 1 - G.half; mask |= 1u; }
    else nb[0] = G.zero;
    // -x
    if (s.i >= 1) { nb[1] = o + p - 1; mask |= 2u; }
    else if (BC == BC_PERIODIC) { nb[1] = o + G.half - 1; mask |= 2u; }
    else nb[1] = G.zero;
    // +y
    if (s.j + 1 < G.nj) { nb[2] = o + G.half; mask |= 4u; }
    else if (BC == BC_PERIODIC) { nb[2] = o - (G.nj - 1) * G.half; mask |= 4u; }
    else nb[2] = G.zero;
    // -y
    if (s.j >= 1) { nb[3] = o - G.half; mask |= 8u; }
    else if (BC == BC_PERIODIC) { nb[3] = o + (G.nj - 1) * G.half; mask |= 8u; }
    else nb[3] = G.zero;
    // +z
    if (s.k + 1 < G.nk) { nb[4] = o + G.plane; mask |= 16u; }
    else if (BC == BC_PERIODIC) { nb[4] = o - (G.nk - 1) * G.plane; mask |= 16u; }
    else nb[4] = G.zero;
    // -z
    if (s.k >= 1) { nb[5] = o - G.plane; mask |= 32u; }
    else if (BC == BC_PERIODIC) { nb[5] = o + (G.nk - 1) * G.plane; mask |= 32u; }
    else nb[5] = G.zero;
    return mask;
}

// Block order for banded launches.  Workgroups go round-robin to the 8 XCDs
// (workgroup b to XCD b % 8), each with its own L2.  With band = P > 0 the
// lattice's blocks are taken in groups of P (one plane's blocks), each group
// cut into 8 bands of ceil(P / 8) blocks (rows), and XCD x takes band x of
// every group, groups in order: a row's y neighbours then sit in the same L2
// except at the band edges, and the z neighbours of a band (P blocks later)
// stay in that L2 while its next planes run.  When P is not a multiple of 8
// the last band of a group is short and the grid holds 8 ceil(P / 8) blocks
// per group; the padding blocks get kIdleBlock and do no work.  band = 0
// keeps the identity order.
constexpr uint32_t kIdleBlock = 0xffffffffu;
__device__ inline uint32_t band_block(uint32_t b, uint32_t band)
{
    if (band == 0u) return b;
    const uint32_t per = (band + 7u) >> 3;
    const uint32_t x = b & 7u, n = b >> 3;
    const uint32_t p = n / per, r = n - p * per, q = x * per + r;
    return q < band ? p * band + q : kIdleBlock;
}

// Block of a sweep launch: bits 0..30 of `band` as band_block, bit 31 = walk
// the grid last to first (a launch then starts on the planes the launch before
// it finished with, which the memory-side cache still holds; a banded grid is
// a multiple of 8 blocks, so reversed block b runs on XCD 7 - b % 8 and keeps
// the bands).
__device__ inline uint32_t sweep_block(uint32_t band)
{
    const uint32_t b = (band >> 31) ? gridDim.x - 1u - blockIdx.x : blockIdx.x;
    return band_block(b, band & 0x7fffffffu);
}

// Lane l - 1's (DPP wave_shr:1) and lane l + 1's (wave_shl:1) copy of a
// value, in one VALU move per 32 bits; lane 0 / lane 63 get zero (bound_ctrl),
// which the callers never use, so no destination needs clearing first.
__device__ inline uint32_t lane_prev(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xf, 0xf, true);
}
__device__ inline uint32_t lane_next(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xf, 0xf, true);
}
__device__ inline double lane_prev(double v)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    return __longlong_as_double((long long)(((uint64_t)lane_prev((uint32_t)(u >> 32)) << 32) | lane_prev((uint32_t)u)));
}
__device__ inline double lane_next(double v)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    return __longlong_as_double((long long)(((uint64_t)lane_next((uint32_t)(u >> 32)) << 32) | lane_next((uint32_t)u)));
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
struct U4 {
    uint32_t x, y, z, w;
};

// Each round's two 32x32->64 products are written as 64-bit multiplies so
// they lower to one v_mad_u64_u32 each instead of a v_mul_hi_u32 /
// v_mul_lo_u32 pair.
// a ^ b ^ c; on the device one v_bitop3_b32 (truth table 0x96) instead of two
// v_xor_b32.
__host__ __device__ inline uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        // round 0 mixes counter words that are often wave-uniform (scalar
        // instructions); from round 1 on every word varies per lane
        if (r == 0)
            c = U4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
        else
            c = U4{xor3((uint32_t)(p1 >> 32), c.y, k0), (uint32_t)p1, xor3((uint32_t)(p0 >> 32), c.w, k1),
                   (uint32_t)p0};
    }
    return c;
}

// 53-bit uniform in [0,1) from two words (hi first): X = (hi:lo) >> 11 is
// A 2^32 + B with A = hi >> 11 (21 bits) and B the next 32 bits, so
// X 2^-53 = A 2^-21 + B 2^-53, both terms and their sum exact in binary64
// (two 32-bit conversions and one fma instead of a 64-bit conversion).
__host__ __device__ inline double u53(uint32_t hi, uint32_t lo)
{
    const uint32_t a = hi >> 11, b = (hi << 21) | (lo >> 11);
    return fma((double)a, 0x1.0p-21, (double)b * 0x1.0p-53);
}

// Deterministic-mode stream u_n: block n/2 with counter (q_lo, q_hi, 0,
// 0x5354524D) and key (seed_lo, seed_hi); even n -> words (y, x), odd n ->
// words (w, z).
__host__ __device__ inline double stream_uniform(uint64_t seed, uint64_t n)
{
    uint64_t q = n >> 1;
    U4 w = philox4x32_10(U4{(uint32_t)q, (uint32_t)(q >> 32), 0u, 0x5354524Du}, (uint32_t)seed,
                         (uint32_t)(seed >> 32));
    return (n & 1) ? u53(w.w, w.z) : u53(w.y, w.x);
}

// Production site-keyed draws: one block per site, sweep and colour.  Counter
// (m, istep, (run_index << 4) | (colour << 1), replica), key (seed_lo,
// seed_hi).  Director trial of sweep t: block t, u_phi = x 2^-32 (32 bits),
// u_acc = (y, z >> 16) as a 48-bit fraction.  Chirality trial of sweep t:
// u_acc = (w, z & 0xffff) as a 48-bit fraction of block spin_sweep(colour, t)
// -- block t for colour 1, block t + 1 for colour 0 -- so that each fused
// launch (colour 1: both trials of sweep t; colour 0: the chirality trial of
// sweep t and the director trial of sweep t + 1) evaluates one block per site.
__host__ __device__ inline U4 site_block(uint64_t seed, uint32_t run_index, uint32_t replica, uint32_t m,
                                uint32_t istep, int colour)
{
    return philox4x32_10(U4{m, istep, (run_index << 4) | ((uint32_t)colour << 1), replica}, (uint32_t)seed,
                         (uint32_t)(seed >> 32));
}

// hi 2^-32 + lo16 2^-48 (exact: one conversion each and one fma).
__host__ __device__ inline double u48(uint32_t hi, uint32_t lo16)
{
    return fma((double)hi, 0x1.0p-32, (double)lo16 * 0x1.0p-48);
}
__host__ __device__ inline uint32_t spin_sweep(int colour, uint32_t istep) { return colour ? istep : istep + 1u; }
__host__ __device__ inline double site_u_phi(const U4 &w) { return (double)w.x * 0x1.0p-32; }
__host__ __device__ inline double site_u_dir(const U4 &w) { return u48(w.y, w.z >> 16); }
__host__ __device__ inline double site_u_spin(const U4 &w) { return u48(w.w, w.z & 0xffffu); }

// ---------------------------------------------------------------------------
__device__ inline double fl32(double x) { return (double)(float)x; }

// Metropolis test u <= exp(x) for x <= 0, with the same result as evaluating
// exp(x) in binary64 but usually without doing so.  e = 2^RN32(RN32(x) *
// RN32(log2 e)) from the hardware exp2 (v_exp_f32, 1 ulp) has
//   |e / exp(x) - 1| <= |x| 1.5 2^-23 + 2^-23 + O(2^-46)
// for -80 <= x <= 0 (argument rounding |x| 2^-24, constant and product
// rounding |x| log2e 1.5 2^-24, times ln 2 when moved into the result).  The
// bracket below uses r = |x| 2^-20 + 2^-17, over 5x that bound, so u < e(1-r)
// and u > e(1+r) decide the test; u inside the bracket (probability ~2 e r)
// takes the binary64 exp.  For x < -80 the bracket at -80 bounds exp(x) from
// above, so only a rejection is decided there.
__device__ inline void exp_bracket(double x, double &lo, double &hi)
{
    const float xf = (float)fmax(x, -80.0);
    const double e = (double)__builtin_amdgcn_exp2f(xf * 1.44269504f);
    const double r = fma(-(double)xf, 0x1.0p-20, 0x1.0p-17);
    lo = x >= -80.0 ? e * (1.0 - r) : 0.0;
    hi = e * (1.0 + r);
}

__device__ inline bool accept_exp(double x, double u)
{
    double lo, hi;
    exp_bracket(x, lo, hi);
    if (u < lo) return true;
    if (u > hi) return false;
    return u <= exp(x);
}

// x / b from y = RN(1/b): q = RN(x y), r = x - b q (exact with an FMA),
// q' = RN(q + r y) = RN(x / b) (Markstein, IBM J. Res. Dev. 34(1), 1990, for
// binary64 without underflow or overflow).  Three FMA-unit operations replace
// a full IEEE division when several quotients share a divisor.
__host__ __device__ inline double div_by_recip(double x, double b, double y)
{
    const double q = x * y;
    const double r = fma(-q, b, x);
    return fma(r, y, q);
}

// Neighbour data gathered once per trial.
// Neighbour data gathered once per trial into registers.  The energy
// functions below take any neighbour accessor with get(b, x, y, z), spin(b)
// and mask; the k-march kernels pass one that reads LDS at every use.
struct Nbr {
    double x[6], y[6], z[6];
    int s[6];
    uint32_t mask;
    __device__ void get(int b, double &mx, double &my, double &mz) const
    {
        mx = x[b];
        my = y[b];
        mz = z[b];
    }
    __device__ int spin(int b) const { return s[b]; }
};

// Reference-compatible site energy E(n, snew) (chll/chll_mod.f90:363-412):
// bonds in the order +x, -x, +y, -y, +z, -z; each bond
//   e = fl32( (e + (1 - c*c)) -/+ ((K*c)*q)*f ),  c = (nx*mx + ny*my) + nz*mz,
//   f = 0.5 * (snew + s_b);
// then e = fl32(e - g*snew).  Absent bonds leave e unchanged.
template <class NB>
__device__ inline double energy_compat(const NB &N, double nx, double ny, double nz, int snew, double g, double KK)
{
    double e = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        double mx, my, mz;
        N.get(b, mx, my, mz);
        const double dott = nx * mx + ny * my + nz * mz;
        double cr;
        if (b < 2) cr = ny * mz - nz * my;
        else if (b < 4) cr = nz * mx - nx * mz;
        else cr = nx * my - ny * mx;
        const double sfac = 0.5 * (double)(snew + N.spin(b));
        // an absent bond's neighbour reads zero (c = cr = 0), so with 0 in
        // place of the 1 its term is +-0 and e stays as it is: the select
        // moves from each sum to the constant, which the site's two energies
        // share
        const double one = (N.mask >> b) & 1u ? 1.0 : 0.0;
        double r;
        if ((b & 1) == 0) r = e + (one - dott * dott) - KK * dott * cr * sfac;
        else r = e + (one - dott * dott) + KK * dott * cr * sfac;
        e = fl32(r);
    }
    return fl32(e - g * (double)snew);
}

// ---------------------------------------------------------------------------
// Pure binary64 variants.  Per bond b with neighbour director m_b:
//   c_b = n.m_b, q_b = (n x m_b)_axis(b), sgn_b = +1 for +x/+y/+z, -1 otherwise.
// Director move n -> n' at fixed spin:
//   dE = sum_b [ (c_b^2 - c'_b^2) - sgn_b K f_b (c'_b q'_b - c_b q_b) ].
// Chirality flip sigma -> -sigma at fixed director:
//   dE = sigma * (K * sum_b sgn_b c_b q_b + 2 g).
template <class NB>
__device__ inline double cq_sum(const NB &N, double nx, double ny, double nz)
{
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        double mx, my, mz;
        N.get(b, mx, my, mz);
        const double c = fma(nz, mz, fma(ny, my, nx * mx));
        double q;
        if (b < 2) q = fma(ny, mz, -nz * my);
        else if (b < 4) q = fma(nz, mx, -nx * mz);
        else q = fma(nx, my, -ny * mx);
        acc = (b & 1) == 0 ? fma(c, q, acc) : fma(-c, q, acc);
    }
    return acc;
}

// Weighted sum for the director move: sum_b sgn_b f_b c_b q_b (f_b = (sigma +
// s_b)/2) and sum_b c_b^2.  Absent bonds read the zero site (c = q = 0) and add
// exactly nothing.
template <class NB>
__device__ inline double cqf_sum(const NB &N, int sigma, double nx, double ny, double nz, double &csq)
{
    double acc = 0.0, c2 = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        double mx, my, mz;
        N.get(b, mx, my, mz);
        const double fb = (b & 1) ? -0.5 * (double)(sigma + N.spin(b)) : 0.5 * (double)(sigma + N.spin(b));
        const double c = fma(nz, mz, fma(ny, my, nx * mx));
        double q;
        if (b < 2) q = fma(ny, mz, -nz * my);
        else if (b < 4) q = fma(nz, mx, -nx * mz);
        else q = fma(nx, my, -ny * mx);
        acc = fma(c * q, fb, acc);
        c2 = fma(c, c, c2);
    }
    csq = c2;
    return acc;
}

// cqf_sum and cq_sum of one director in one pass over its bonds: returns the
// weighted director sum (cqf_sum), the c^2 sum in csq and the chirality sum
// sum_b sgn_b c_b q_b (cq_sum) in cq.  Each sum takes the same operands in the
// same order as the separate functions, so the results are bit-identical.
template <class NB>
__device__ inline double cqf_cq_sum(const NB &N, int sigma, double nx, double ny, double nz, double &csq, double &cq)
{
    double acc = 0.0, c2 = 0.0, a = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        double mx, my, mz;
        N.get(b, mx, my, mz);
        const double fb = (b & 1) ? -0.5 * (double)(sigma + N.spin(b)) : 0.5 * (double)(sigma + N.spin(b));
        const double c = fma(nz, mz, fma(ny, my, nx * mx));
        double q;
        if (b < 2) q = fma(ny, mz, -nz * my);
        else if (b < 4) q = fma(nz, mx, -nx * mz);
        else q = fma(nx, my, -ny * mx);
        acc = fma(c * q, fb, acc);
        c2 = fma(c, c, c2);
        a = (b & 1) == 0 ? fma(c, q, a) : fma(-c, q, a);
    }
    csq = c2;
    cq = a;
    return acc;
}

// The chirality sum of director n (cq_sum, returned), its c^2 sum and the
// bond products c_b q_b, which weight_cq turns into cqf_sum for any spin.
// Same operands and order as the separate functions: bit-identical sums.
template <class NB>
__device__ inline double cq_products(const NB &N, double nx, double ny, double nz, double &csq, double cqb[6])
{
    double c2 = 0.0, a = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        double mx, my, mz;
        N.get(b, mx, my, mz);
        const double c = fma(nz, mz, fma(ny, my, nx * mx));
        double q;
        if (b < 2) q = fma(ny, mz, -nz * my);
        else if (b < 4) q = fma(nz, mx, -nx * mz);
        else q = fma(nx, my, -ny * mx);
        cqb[b] = c * q;
        c2 = fma(c, c, c2);
        a = (b & 1) == 0 ? fma(c, q, a) : fma(-c, q, a);
    }
    csq = c2;
    return a;
}

template <class NB>
__device__ inline double weight_cq(const NB &N, int sigma, const double cqb[6])
{
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const double fb = (b & 1) ? -0.5 * (double)(sigma + N.spin(b)) : 0.5 * (double)(sigma + N.spin(b));
        acc = fma(cqb[b], fb, acc);
    }
    return acc;
}

// cqf_sum_mag with, for a later chirality trial's screen, the chirality sum
// sum_b sgn_b c_b q_b in cq (cq_sum's operands and order) and sum_b |c_b q_b|
// in abscq, in the same pass over the bonds.
template <class NB>
__device__ inline double cqf_sum_mag_cq(const NB &N, int sigma, double nx, double ny, double nz, double &csq,
                                        double &mag, double &cq, double &abscq)
{
    double acc = 0.0, c2 = 0.0, m = 0.0, a = 0.0, ab = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        double mx, my, mz;
        N.get(b, mx, my, mz);
        const double fb = (b & 1) ? -0.5 * (double)(sigma + N.spin(b)) : 0.5 * (double)(sigma + N.spin(b));
        const double c = fma(nz, mz, fma(ny, my, nx * mx));
        double q;
        if (b < 2) q = fma(ny, mz, -nz * my);
        else if (b < 4) q = fma(nz, mx, -nx * mz);
        else q = fma(nx, my, -ny * mx);
        const double p = c * q;
        acc = fma(p, fb, acc);
        c2 = fma(c, c, c2);
        m = fma(fabs(p), fabs(fb), m);
        a = (b & 1) == 0 ? fma(c, q, a) : fma(-c, q, a);
        ab += fabs(p);
    }
    csq = c2;
    mag = m;
    cq = a;
    abscq = ab;
    return acc;
}

// cq_products with the chirality screen's magnitude sum_b |c_b q_b|
// max(|sigma|, |s_b|) in w; weight_cq_mag then gives cqf_sum of any spin with
// its magnitude sum_b |c_b q_b f_b|.
template <class NB>
__device__ inline double cq_products_mag(const NB &N, int sigma, double nx, double ny, double nz, double &csq,
                                         double cqb[6], double &w)
{
    const double sa = fabs((double)sigma);
    double c2 = 0.0, a = 0.0, ww = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        double mx, my, mz;
        N.get(b, mx, my, mz);
        const double c = fma(nz, mz, fma(ny, my, nx * mx));
        double q;
        if (b < 2) q = fma(ny, mz, -nz * my);
        else if (b < 4) q = fma(nz, mx, -nx * mz);
        else q = fma(nx, my, -ny * mx);
        cqb[b] = c * q;
        c2 = fma(c, c, c2);
        a = (b & 1) == 0 ? fma(c, q, a) : fma(-c, q, a);
        ww = fma(fabs(cqb[b]), fmax(sa, fabs((double)N.spin(b))), ww);
    }
    csq = c2;
    w = ww;
    return a;
}

template <class NB>
__device__ inline double weight_cq_mag(const NB &N, int sigma, const double cqb[6], double &mag)
{
    double acc = 0.0, m = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const double fb = (b & 1) ? -0.5 * (double)(sigma + N.spin(b)) : 0.5 * (double)(sigma + N.spin(b));
        acc = fma(cqb[b], fb, acc);
        m = fma(fabs(cqb[b]), fabs(fb), m);
    }
    mag = m;
    return acc;
}

// ---------------------------------------------------------------------------
// Screen for the reference-compatible decisions.  energy_compat rounds its
// running sum to binary32 at each of its seven steps (six bonds, the field
// term): r_k = e_{k-1} + t_k, e_k = fl32(r_k), |e_k - r_k| <= 2^-24 |r_k|.
// Every |r_k| is at most A + |g s|, A = sum_b |t_b| <= sum_b (1 + c_b^2 +
// |K c_b q_b f_b|), so one site energy lies within 7 2^-24 (A + |g s|) of the
// exact one, and the binary64 operations (the terms, the difference of the
// two rounded energies, the direct binary64 difference dE) add below 2^-40
// of it.  Summed over a trial's two energies as M, the reference's energy
// difference dE_c is within B = 2^-21 M + 2^-120 of dE (8 2^-24 covers
// 7 2^-24 and the binary64 terms; 2^-120 covers binary32 underflow).
//
// The reference's decision, accept iff dE_c < 0 or u <= exp(RN(-dE_c / kT)),
// is monotone in dE_c for kT > 0: it accepts every dE_c below some threshold.
// With del = B/kT <= 1/4 and x = -dE/kT bracketed by exp_bracket as
// lo <= exp(x) <= hi, every admissible dE_c has
//   exp(x) (1 - del) <= exp(-dE_c / kT) <= exp(x) (1 + 2 del),
// so u < lo (1 - del) accepts and u > hi (1 + 2 del) rejects, whatever dE_c
// is (the brackets' 5x margin also covers the few ulp of x and of the
// binary64 exp).  Returns 1 (accepted), 0 (rejected) or -1 (not decided:
// the caller evaluates the two rounded energies).  scale = 0 turns the screen
// off; the library never passes a scale below 2^-21.  The probability of -1
// is about 2 (del + 2^-17): below 1e-4 per trial at kT >= 0.2 and K, g of
// order one.
__device__ inline int compat_screen(double dE, double mag, double scale, double kbt, double inv_kbt, double u)
{
    if (!(scale > 0.0 && kbt > 0.0 && mag < 0x1.0p100)) return -1;
    const double B = fma(scale, mag, 0x1.0p-120);
    if (dE + B < 0.0) return 1;
    const double del = B * inv_kbt;
    if (!(del <= 0.25)) return -1;
    double lo, hi;
    exp_bracket(-dE * inv_kbt, lo, hi);
    if (u < lo * (1.0 - del)) return 1;
    if (u > hi * fma(2.0, del, 1.0)) return 0;
    return -1;
}

// cqf_sum with the screen's magnitude: also returns sum_b |c_b q_b f_b| in mag.
template <class NB>
__device__ inline double cqf_sum_mag(const NB &N, int sigma, double nx, double ny, double nz, double &csq, double &mag)
{
    double acc = 0.0, c2 = 0.0, a = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        double mx, my, mz;
        N.get(b, mx, my, mz);
        const double fb = (b & 1) ? -0.5 * (double)(sigma + N.spin(b)) : 0.5 * (double)(sigma + N.spin(b));
        const double c = fma(nz, mz, fma(ny, my, nx * mx));
        double q;
        if (b < 2) q = fma(ny, mz, -nz * my);
        else if (b < 4) q = fma(nz, mx, -nx * mz);
        else q = fma(nx, my, -ny * mx);
        const double cq = c * q;
        acc = fma(cq, fb, acc);
        c2 = fma(c, c, c2);
        a = fma(fabs(cq), fabs(fb), a);
    }
    csq = c2;
    mag = a;
    return acc;
}

// cq_sum with the screen's magnitudes of a chirality flip sigma -> -sigma:
// the c^2 sum in csq and sum_b |c_b q_b| max(|sigma|, |s_b|) in mag, which is
// sum_b |c_b q_b| (|f_b(sigma)| + |f_b(-sigma)|).
template <class NB>
__device__ inline double cq_sum_mag(const NB &N, int sigma, double nx, double ny, double nz, double &csq,
                                    double &mag)
{
    const double sa = fabs((double)sigma);
    double acc = 0.0, c2 = 0.0, w = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        double mx, my, mz;
        N.get(b, mx, my, mz);
        const double c = fma(nz, mz, fma(ny, my, nx * mx));
        double q;
        if (b < 2) q = fma(ny, mz, -nz * my);
        else if (b < 4) q = fma(nz, mx, -nx * mz);
        else q = fma(nx, my, -ny * mx);
        acc = (b & 1) == 0 ? fma(c, q, acc) : fma(-c, q, acc);
        c2 = fma(c, c, c2);
        w = fma(fabs(c * q), fmax(sa, fabs((double)N.spin(b))), w);
    }
    csq = c2;
    mag = w;
    return acc;
}

// a * b + c rounded once, with c taken from a scalar register pair: for a
// constant c this is one v_fma_f64 fed by two scalar moves, where the compiler
// otherwise rebuilds c in vector registers (two v_mov_b32) for a v_fmac_f64.
__device__ inline double fma_sc(double a, double b, double c)
{
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

// sin and cos of 2 pi u for u in [0, 1): w = 4u and f = w - rint(w) are exact,
// so 2 pi u = n pi/2 + x with x = f pi/2 in [-pi/4, pi/4]; sin x and cos x
// from their Taylor series through x^17 and x^18 (truncation below 1e-19 on
// that interval), then rotated by the quadrant n.
__device__ inline void sincos_2pi(double u, double &s, double &c)
{
    const double w = 4.0 * u;
    const double n = rint(w);
    const double x = (w - n) * 1.5707963267948966;
    const double z = x * x;
    double ps = -1.0 / 355687428096000.0;
    ps = fma_sc(ps, z, 1.0 / 1307674368000.0);
    ps = fma_sc(ps, z, -1.0 / 6227020800.0);
    ps = fma_sc(ps, z, 1.0 / 39916800.0);
    ps = fma_sc(ps, z, -1.0 / 362880.0);
    ps = fma_sc(ps, z, 1.0 / 5040.0);
    ps = fma_sc(ps, z, -1.0 / 120.0);
    ps = fma_sc(ps, z, 1.0 / 6.0);
    const double sx = fma(-x * z, ps, x);
    double pc = 1.0 / 6402373705728000.0;
    pc = fma_sc(pc, z, -1.0 / 20922789888000.0);
    pc = fma_sc(pc, z, 1.0 / 87178291200.0);
    pc = fma_sc(pc, z, -1.0 / 479001600.0);
    pc = fma_sc(pc, z, 1.0 / 3628800.0);
    pc = fma_sc(pc, z, -1.0 / 40320.0);
    pc = fma_sc(pc, z, 1.0 / 720.0);
    pc = fma_sc(pc, z, -1.0 / 24.0);
    pc = fma(pc, z, 0.5);
    const double cx = fma(-z, pc, 1.0);
    const int q = (int)n;
    const bool swap = q & 1;
    double ss = swap ? cx : sx;
    double cc = swap ? sx : cx;
    s = (q & 2) ? -ss : ss;
    c = ((q + 1) & 2) ? -cc : cc;
}

// 1/sqrt(x) for a normal x > 0: the hardware estimate (v_rsq_f64) refined
// by one step y + y e (1/2 + 3/8 e), e = 1 - x y^2, which leaves an error of
// order e^3 (below the rounding of the result).  No scaling for denormal or
// infinite x, which the cone trial never forms from unit directors.
__device__ inline double rsqrt_normal(double x)
{
    const double y0 = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y0, y0, 1.0);
    return fma(y0 * e, fma(e, 0.375, 0.5), y0);
}

// Cone trial in binary64 with reciprocal square roots: the same geometry
// (tangent basis u = z x n / |.|, v = n x u, displacement d on a cone of
// half-angle atan d at azimuth 2 pi u_phi), without the compat operation order.
// For a unit director |n x u| = |n| = 1, so v = (-nz nx, -nz ny, rho^2) / rho
// shares u's normalisation (rounding leaves |v| within a few ulp of 1; the
// trial vector is renormalised at the end).  Same branch as the reference for
// |nz| > 0.999 (chll/chll_mod.f90:149-157).
__device__ inline void trial_fp64(double nnx, double nny, double nnz, double d, double u_phi, double &ox,
                                  double &oy, double &oz)
{
    double sphi, cphi;
    sincos_2pi(u_phi, sphi, cphi);
    double tx, ty, tz;
    if (__builtin_expect(fabs(nnz) > 0.999, 0)) {
        tx = fma(d, cphi, nnx);
        ty = fma(d, sphi, nny);
        tz = nnz;
    } else {
        const double rho2 = fma(nnx, nnx, nny * nny);
        const double ir = rsqrt_normal(rho2);
        const double ux = -nny * ir, uy = nnx * ir;
        const double nzr = nnz * ir;
        const double vx = -nzr * nnx, vy = -nzr * nny, vz = rho2 * ir;
        const double dc = d * cphi, ds = d * sphi;
        tx = fma(ds, vx, fma(dc, ux, nnx));
        ty = fma(ds, vy, fma(dc, uy, nny));
        tz = fma(ds, vz, nnz);
    }
    const double inv = rsqrt_normal(fma(tz, tz, fma(ty, ty, tx * tx)));
    ox = tx * inv; oy = ty * inv; oz = tz * inv;
}

}  // namespace chll
