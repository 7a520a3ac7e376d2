// chll_grad.hip -- lattice energy and its gradient (gfx950).
//
// E(n, s; eps, K, g) = sum over bonds (a, b = a + e_axis) of
//     eps (1 - c^2) - K c q f,   c = n_a . n_b,  q = (n_a x n_b)_axis,
//     f = (s_a + s_b) / 2,
// minus g sum_a s_a -- the Hamiltonian the sweep samples (chll/chll_mod.f90:
// 369-411 with each bond counted once; the reference has eps = 1).  Open
// boundaries drop the bonds that leave the lattice, periodic ones wrap.
//
// One thread per site (two adjacent sites of a row when ni is even) gathers
// the bond terms of its +x, +y, +z bonds (each bond once, for E and the
// parameter derivatives) and the derivative of all six of its bonds with
// respect to its own director:
//   d t / d n_a = -2 eps c n_b - K f (q n_b + c dq/dn_a)       (a lower end)
//   d t / d n_b = -2 eps c n_a - K f (q n_a + c dq/dn_b)       (b upper end)
// with, for axis x, dq/dn_a = (0, n_bz, -n_by), dq/dn_b = (0, -n_az, n_ay) and
// cyclically for y, z.  The gradient returned is the tangent projection
// (I - n_a n_a^T) dE/dn_a.  No atomics: each block writes partial sums that a
// second kernel reduces in a fixed order, so results are reproducible.
//
// Layout: plain Fortran order per lattice, field c of lattice b at
// n + (b*3 + c)*N (i fastest), spins int8 at s + b*N.
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/chll_mod.h"
#include "chll_device.h"

namespace chll {

constexpr int kGradBlock = 256;

struct GradArgs {
    const double *n;
    const int8_t *s;
    const double *K, *g, *eps;  // per lattice
    double *dEdn;               // B x 3 x N, or null
    double *part;               // B x nblk x 4: sum eps-term (1-c^2), sum c q f, sum s, unused
    uint32_t ni, nj, nk, N, nblk;
    uint32_t band;              // XCD row bands (band_block), 0 = lattice order
    FastDiv fd_ni, fd_nj;
    int bc;
};

__device__ inline void dq_axis(int axis, double ax, double ay, double az, double bx, double by, double bz, double &q,
                               double da[3], double db[3])
{
    if (axis == 0) {
        q = fma(ay, bz, -az * by);
        da[0] = 0.0; da[1] = bz; da[2] = -by;
        db[0] = 0.0; db[1] = -az; db[2] = ay;
    } else if (axis == 1) {
        q = fma(az, bx, -ax * bz);
        da[0] = -bz; da[1] = 0.0; da[2] = bx;
        db[0] = az; db[1] = 0.0; db[2] = -ax;
    } else {
        q = fma(ax, by, -ay * bx);
        da[0] = by; da[1] = -bx; da[2] = 0.0;
        db[0] = -ay; db[1] = ax; db[2] = 0.0;
    }
}

// g += w m + kf (q m + c dq) with fused multiply-adds; components whose dq is
// identically zero reduce to products.
__device__ inline void grad_acc(double &gx, double &gy, double &gz, double w, double kf, double q, double c,
                                double mx, double my, double mz, const double dq[3])
{
    gx = fma(kf, fma(c, dq[0], q * mx), fma(w, mx, gx));
    gy = fma(kf, fma(c, dq[1], q * my), fma(w, my, gy));
    gz = fma(kf, fma(c, dq[2], q * mz), fma(w, mz, gz));
}

template <int BC>
__global__ __launch_bounds__(kGradBlock) void k_energy_grad(GradArgs A)
{
    const uint32_t b = blockIdx.y;
    const uint32_t l = band_block(blockIdx.x, A.band) * kGradBlock + threadIdx.x;
    const double *nx = A.n + (size_t)b * 3 * A.N, *ny = nx + A.N, *nz = ny + A.N;
    const int8_t *s = A.s + (size_t)b * A.N;
    const double K = A.K[b], m2eps = -2.0 * A.eps[b];
    double sum_e = 0.0, sum_cqf = 0.0, sum_s = 0.0;
    if (l < A.N) {
        const uint32_t row = fdiv(l, A.fd_ni);
        const uint32_t i = l - row * A.ni;
        const uint32_t k = fdiv(row, A.fd_nj);
        const uint32_t j = row - k * A.nj;
        const double ax = nx[l], ay = ny[l], az = nz[l];
        const int sa = s[l];
        double gx = 0.0, gy = 0.0, gz = 0.0;
        const uint32_t co[3] = {i, j, k};
        const uint32_t dim[3] = {A.ni, A.nj, A.nk};
        const uint32_t stride[3] = {1u, A.ni, A.ni * A.nj};
#pragma unroll
        for (int axis = 0; axis < 3; ++axis) {
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {  // dir 0: bond to the + neighbour (a lower), 1: to the - neighbour
                uint32_t q;
                bool ok;
                if (dir == 0) {
                    ok = co[axis] + 1 < dim[axis];
                    q = ok ? l + stride[axis] : (BC == BC_PERIODIC ? l - (dim[axis] - 1) * stride[axis] : l);
                    ok = ok || (BC == BC_PERIODIC);
                } else {
                    ok = co[axis] >= 1;
                    q = ok ? l - stride[axis] : (BC == BC_PERIODIC ? l + (dim[axis] - 1) * stride[axis] : l);
                    ok = ok || (BC == BC_PERIODIC);
                }
                if (!ok) continue;
                const double bx = nx[q], by = ny[q], bz = nz[q];
                const double f = 0.5 * (double)(sa + s[q]);
                double qv, da[3], db[3];
                double c;
                if (dir == 0) {
                    c = fma(az, bz, fma(ay, by, ax * bx));
                    dq_axis(axis, ax, ay, az, bx, by, bz, qv, da, db);
                    sum_e += fma(-c, c, 1.0);
                    sum_cqf = fma(c * qv, f, sum_cqf);
                    grad_acc(gx, gy, gz, m2eps * c, -K * f, qv, c, bx, by, bz, da);
                } else {
                    // the bond (q, l): q is the lower end, l the upper end
                    c = fma(az, bz, fma(ay, by, ax * bx));
                    dq_axis(axis, bx, by, bz, ax, ay, az, qv, da, db);
                    grad_acc(gx, gy, gz, m2eps * c, -K * f, qv, c, bx, by, bz, db);
                }
            }
        }
        sum_s = (double)sa;
        if (A.dEdn) {
            const double p = gx * ax + gy * ay + gz * az;
            double *o = A.dEdn + (size_t)b * 3 * A.N;
            o[l] = gx - p * ax;
            o[A.N + l] = gy - p * ay;
            o[2 * (size_t)A.N + l] = gz - p * az;
        }
    }
    // deterministic block sums
    __shared__ double red[kGradBlock / 64][3];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sum_e += __shfl_xor(sum_e, off, 64);
        sum_cqf += __shfl_xor(sum_cqf, off, 64);
        sum_s += __shfl_xor(sum_s, off, 64);
    }
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (lane == 0) { red[wid][0] = sum_e; red[wid][1] = sum_cqf; red[wid][2] = sum_s; }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < kGradBlock / 64; ++w) t += red[w][threadIdx.x];
        A.part[((size_t)b * A.nblk + blockIdx.x) * 4 + threadIdx.x] = t;
    }
}

// Bond (lo, up = lo + e_axis) with spin factor f: adds its energy terms when
// COUNT, its derivative with respect to lo into glo when GLO, and with respect
// to up into gup when GUP.
template <bool COUNT, bool GLO, bool GUP>
__device__ inline void bond(int axis, double lx, double ly, double lz, double ux, double uy, double uz, double f,
                            double m2eps, double K, double &se, double &scqf, double glo[3], double gup[3])
{
    double q, dl[3], du[3];
    const double c = fma(lz, uz, fma(ly, uy, lx * ux));
    dq_axis(axis, lx, ly, lz, ux, uy, uz, q, dl, du);
    if (COUNT) {
        se += fma(-c, c, 1.0);
        scqf = fma(c * q, f, scqf);
    }
    const double w = m2eps * c, kf = -K * f;
    if (GLO) grad_acc(glo[0], glo[1], glo[2], w, kf, q, c, ux, uy, uz, dl);
    if (GUP) grad_acc(gup[0], gup[1], gup[2], w, kf, q, c, lx, ly, lz, du);
}

__device__ inline double2 ld2(const double *p) { return *reinterpret_cast<const double2 *>(p); }
__device__ inline void spin_pair(const int8_t *s, uint32_t q, int &lo, int &hi)
{
    const uint32_t v = *reinterpret_cast<const uint16_t *>(s + q);
    lo = (int)(int8_t)(v & 0xffu);
    hi = (int)(int8_t)(v >> 8);
}

// Two sites (l, l + 1), l even, per thread when ni is even: both lie in one
// row, so their directors, their +-y and +-z neighbours and their gradient
// move as 16-byte words and their spins as 2-byte words; the x bond between
// them is evaluated once for both ends.  Requires 16-byte aligned director
// and gradient arrays and 2-byte aligned spins (checked on the host).
template <int BC>
__global__ __launch_bounds__(kGradBlock) void k_energy_grad2(GradArgs A)
{
    const uint32_t b = blockIdx.y;
    const uint32_t l = 2u * (band_block(blockIdx.x, A.band) * kGradBlock + threadIdx.x);
    const double *nx = A.n + (size_t)b * 3 * A.N, *ny = nx + A.N, *nz = ny + A.N;
    const int8_t *s = A.s + (size_t)b * A.N;
    const double K = A.K[b], m2eps = -2.0 * A.eps[b];
    double sum_e = 0.0, sum_cqf = 0.0, sum_s = 0.0;
    if (l < A.N) {
        const uint32_t row = fdiv(l, A.fd_ni);
        const uint32_t i = l - row * A.ni;
        const uint32_t k = fdiv(row, A.fd_nj);
        const uint32_t j = row - k * A.nj;
        const uint32_t plane = A.ni * A.nj;
        const double2 vx = ld2(nx + l), vy = ld2(ny + l), vz = ld2(nz + l);
        const uint32_t sw = *reinterpret_cast<const uint16_t *>(s + l);
        const int sa = (int)(int8_t)(sw & 0xffu), sb = (int)(int8_t)(sw >> 8);
        // Consecutive lanes hold consecutive pairs, so inside a row the site
        // after the pair (l + 2) is the next lane's first site and the site
        // before it (l - 1) the previous lane's second: they move between
        // lanes by DPP, and only the edge lanes of a wave and periodic wraps
        // load them.
        const uint32_t lane = threadIdx.x & 63u;
        const double nxx = lane_next(vx.x), nxy = lane_next(vy.x), nxz = lane_next(vz.x);
        const uint32_t nsw = lane_next(sw);
        const double pxx = lane_prev(vx.y), pxy = lane_prev(vy.y), pxz = lane_prev(vz.y);
        const uint32_t psw = lane_prev(sw);
        double ga[3] = {0.0, 0.0, 0.0}, gb[3] = {0.0, 0.0, 0.0};
        // x: a -> b inside the pair, b -> the site after it, the site before a -> a
        bond<true, true, true>(0, vx.x, vy.x, vz.x, vx.y, vy.y, vz.y, 0.5 * (double)(sa + sb), m2eps, K, sum_e,
                               sum_cqf, ga, gb);
        {
            const bool in = i + 2u < A.ni;
            if (in || BC == BC_PERIODIC) {
                double bx, by, bz;
                int sq;
                if (in && lane != 63u) {
                    bx = nxx; by = nxy; bz = nxz; sq = (int)(int8_t)(nsw & 0xffu);
                } else {
                    const uint32_t q = in ? l + 2u : l - i;
                    bx = nx[q]; by = ny[q]; bz = nz[q]; sq = s[q];
                }
                bond<true, true, false>(0, vx.y, vy.y, vz.y, bx, by, bz, 0.5 * (double)(sb + sq), m2eps, K, sum_e,
                                        sum_cqf, gb, gb);
            }
        }
        {
            const bool in = i >= 1u;
            if (in || BC == BC_PERIODIC) {
                double bx, by, bz;
                int sq;
                if (in && lane != 0u) {
                    bx = pxx; by = pxy; bz = pxz; sq = (int)(int8_t)(psw >> 8);
                } else {
                    const uint32_t q = in ? l - 1u : l + A.ni - 1u;
                    bx = nx[q]; by = ny[q]; bz = nz[q]; sq = s[q];
                }
                bond<false, false, true>(0, bx, by, bz, vx.x, vy.x, vz.x, 0.5 * (double)(sq + sa), m2eps, K, sum_e,
                                         sum_cqf, ga, ga);
            }
        }
        // y and z: the pair above (a, b lower ends) and the pair below (a, b upper ends)
        const uint32_t co[2] = {j, k}, dim[2] = {A.nj, A.nk}, stride[2] = {A.ni, plane};
#pragma unroll
        for (int ax = 0; ax < 2; ++ax) {
            const int axis = ax + 1;
            {
                const bool in = co[ax] + 1u < dim[ax];
                if (in || BC == BC_PERIODIC) {
                    const uint32_t q = in ? l + stride[ax] : l - (dim[ax] - 1u) * stride[ax];
                    const double2 ux = ld2(nx + q), uy = ld2(ny + q), uz = ld2(nz + q);
                    int ta, tb;
                    spin_pair(s, q, ta, tb);
                    bond<true, true, false>(axis, vx.x, vy.x, vz.x, ux.x, uy.x, uz.x, 0.5 * (double)(sa + ta), m2eps,
                                            K, sum_e, sum_cqf, ga, ga);
                    bond<true, true, false>(axis, vx.y, vy.y, vz.y, ux.y, uy.y, uz.y, 0.5 * (double)(sb + tb), m2eps,
                                            K, sum_e, sum_cqf, gb, gb);
                }
            }
            {
                const bool in = co[ax] >= 1u;
                if (in || BC == BC_PERIODIC) {
                    const uint32_t q = in ? l - stride[ax] : l + (dim[ax] - 1u) * stride[ax];
                    const double2 ux = ld2(nx + q), uy = ld2(ny + q), uz = ld2(nz + q);
                    int ta, tb;
                    spin_pair(s, q, ta, tb);
                    bond<false, false, true>(axis, ux.x, uy.x, uz.x, vx.x, vy.x, vz.x, 0.5 * (double)(ta + sa), m2eps,
                                             K, sum_e, sum_cqf, ga, ga);
                    bond<false, false, true>(axis, ux.y, uy.y, uz.y, vx.y, vy.y, vz.y, 0.5 * (double)(tb + sb), m2eps,
                                             K, sum_e, sum_cqf, gb, gb);
                }
            }
        }
        sum_s = (double)(sa + sb);
        if (A.dEdn) {
            const double pa = fma(ga[2], vz.x, fma(ga[1], vy.x, ga[0] * vx.x));
            const double pb = fma(gb[2], vz.y, fma(gb[1], vy.y, gb[0] * vx.y));
            double *o = A.dEdn + (size_t)b * 3 * A.N + l;
            *reinterpret_cast<double2 *>(o) = make_double2(fma(-pa, vx.x, ga[0]), fma(-pb, vx.y, gb[0]));
            *reinterpret_cast<double2 *>(o + A.N) = make_double2(fma(-pa, vy.x, ga[1]), fma(-pb, vy.y, gb[1]));
            *reinterpret_cast<double2 *>(o + 2 * (size_t)A.N) =
                make_double2(fma(-pa, vz.x, ga[2]), fma(-pb, vz.y, gb[2]));
        }
    }
    __shared__ double red[kGradBlock / 64][3];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sum_e += __shfl_xor(sum_e, off, 64);
        sum_cqf += __shfl_xor(sum_cqf, off, 64);
        sum_s += __shfl_xor(sum_s, off, 64);
    }
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (lane == 0) { red[wid][0] = sum_e; red[wid][1] = sum_cqf; red[wid][2] = sum_s; }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < kGradBlock / 64; ++w) t += red[w][threadIdx.x];
        A.part[((size_t)b * A.nblk + blockIdx.x) * 4 + threadIdx.x] = t;
    }
}

// A pair of sites (l, l + 1) of one row: directors and spins.
struct GPair {
    double2 x, y, z;
    uint32_t s;
};

__device__ inline GPair load_gpair(const double *nx, const double *ny, const double *nz, const int8_t *s, uint32_t q)
{
    return GPair{ld2(nx + q), ld2(ny + q), ld2(nz + q), *reinterpret_cast<const uint16_t *>(s + q)};
}

__device__ inline int gp_lo(uint32_t v) { return (int)(int8_t)(v & 0xffu); }
__device__ inline int gp_hi(uint32_t v) { return (int)(int8_t)(v >> 8); }

// The two-site energy and gradient marching along k.  Block (lattice
// blockIdx.y, tile t of BS consecutive pairs of a plane, slab sl of ks
// planes): each thread takes one pair column through the slab.  Every step
// loads the pair ahead (+z marching up, -z marching down) and carries its own
// pair and the one behind from the steps before, so the pairs of a plane are
// loaded once per slab.  The +-y pairs of a pair are other threads' own pairs
// of the same plane: each step the block puts its pairs in LDS (two buffers,
// one barrier per step), and only the +-y pairs outside the tile (the rows
// beyond its first and last) are loaded, from L2.  Slabs alternate direction,
// so the planes that adjacent slabs both read are read by both at about the
// same time: the slabs of one XCD are contiguous (workgroup x runs on XCD
// x % 8 when the grid is a multiple of 8) and its blocks run tile by tile
// through all of them.  The gradient goes out with
// streaming stores: it is not read again, so it does not take L2 lines from
// the directors.  Every site's bonds are taken in the same order as in
// k_energy_grad2, so the gradient is the same bit for bit; the partial sums
// are per block (tiles x nslab per lattice).
template <int BC, int BS>
__global__ __launch_bounds__(BS) void k_energy_grad_march(GradArgs A, uint32_t tiles, uint32_t nslab, uint32_t ks,
                                                          int whole)
{
    // whole (a 1-D grid, batch a multiple of 8): XCD q = blockIdx.x % 8 takes
    // lattices q, q + 8, ... whole, so no plane of a lattice is read by two
    // XCDs; else (grid (tiles nslab, batch)) every XCD takes a contiguous
    // range of slabs of every lattice
    uint32_t b, x;
    if (whole) {
        const uint32_t q = blockIdx.x & 7u, n = blockIdx.x >> 3, per = tiles * nslab;
        b = q + 8u * (n / per);
        x = n % per;
    } else {
        b = blockIdx.y;
        x = blockIdx.x;
    }
    uint32_t sl, t;
    if (!whole && (nslab & 7u) == 0u) {
        const uint32_t n = x >> 3, spx = nslab >> 3;
        t = n / spx;
        sl = (x & 7u) * spx + n % spx;
    } else {
        t = x / nslab;
        sl = x % nslab;
    }
    const double *nx = A.n + (size_t)b * 3 * A.N, *ny = nx + A.N, *nz = ny + A.N;
    const int8_t *s = A.s + (size_t)b * A.N;
    const double K = A.K[b], m2eps = -2.0 * A.eps[b];
    const uint32_t plane = A.ni * A.nj, half = A.ni >> 1;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t pq = t * BS + tid;  // pair index in its plane
    const uint32_t lp = 2u * pq;       // site a of the pair, in its plane
    const uint32_t k0 = sl * ks, k1 = k0 + ks < A.nk ? k0 + ks : A.nk;
    const bool act = lp < plane;
    const uint32_t j = act ? fdiv(lp, A.fd_ni) : 0u;
    const uint32_t i = lp - j * A.ni;
    const bool up = (sl & 1u) == 0u;
    const int32_t dk = up ? 1 : -1;
    // the +-y pairs (plane pair index) and whether they lie in this tile
    const bool hyp = act && (j + 1u < A.nj || BC == BC_PERIODIC), hym = act && (j >= 1u || BC == BC_PERIODIC);
    const uint32_t qyp = j + 1u < A.nj ? pq + half : pq - (A.nj - 1u) * half;
    const uint32_t qym = j >= 1u ? pq - half : pq + (A.nj - 1u) * half;
    const uint32_t tb = t * BS;
    const bool yp_in = qyp - tb < (uint32_t)BS, ym_in = qym - tb < (uint32_t)BS;
    __shared__ double2 lx[2][BS], ly[2][BS], lz[2][BS];
    __shared__ uint32_t lsp[2][BS];
    // plane index of the neighbour of plane k in direction d (+1 / -1), or nk
    // when it is absent
    auto zn = [&](uint32_t k, int32_t d) -> uint32_t {
        if (d > 0) return k + 1u < A.nk ? k + 1u : (BC == BC_PERIODIC ? 0u : A.nk);
        return k >= 1u ? k - 1u : (BC == BC_PERIODIC ? A.nk - 1u : A.nk);
    };
    double sum_e = 0.0, sum_cqf = 0.0, sum_s = 0.0;
    const uint32_t kb = up ? k0 : k1 - 1u;  // first plane of the walk
    GPair C{}, P{};  // own pair, and the pair behind the walk
    if (act && k0 < k1) {
        C = load_gpair(nx, ny, nz, s, kb * plane + lp);
        const uint32_t kq = zn(kb, -dk);
        if (kq < A.nk) P = load_gpair(nx, ny, nz, s, kq * plane + lp);
    }
    for (uint32_t step = 0; k0 + step < k1; ++step) {  // same trip count in the whole block
        const uint32_t k = up ? k0 + step : k1 - 1u - step;
        const uint32_t l = k * plane + lp;
        const uint32_t buf = step & 1u;
        GPair N{}, Yp{}, Ym{};
        if (act) {
            const uint32_t ka = zn(k, dk);
            if (ka < A.nk) N = load_gpair(nx, ny, nz, s, ka * plane + lp);
            if (hyp && !yp_in) Yp = load_gpair(nx, ny, nz, s, k * plane + 2u * qyp);
            if (hym && !ym_in) Ym = load_gpair(nx, ny, nz, s, k * plane + 2u * qym);
        }
        lx[buf][tid] = C.x; ly[buf][tid] = C.y; lz[buf][tid] = C.z; lsp[buf][tid] = C.s;
        __syncthreads();
        if (act) {
            if (hyp && yp_in) {
                const uint32_t u = qyp - tb;
                Yp = GPair{lx[buf][u], ly[buf][u], lz[buf][u], lsp[buf][u]};
            }
            if (hym && ym_in) {
                const uint32_t u = qym - tb;
                Ym = GPair{lx[buf][u], ly[buf][u], lz[buf][u], lsp[buf][u]};
            }
        }
        const double2 vx = C.x, vy = C.y, vz = C.z;
        const int sa = gp_lo(C.s), sb = gp_hi(C.s);
        const double nxx = lane_next(vx.x), nxy = lane_next(vy.x), nxz = lane_next(vz.x);
        const uint32_t nsw = lane_next(C.s);
        const double pxx = lane_prev(vx.y), pxy = lane_prev(vy.y), pxz = lane_prev(vz.y);
        const uint32_t psw = lane_prev(C.s);
        if (act) {
            double ga[3] = {0.0, 0.0, 0.0}, gb[3] = {0.0, 0.0, 0.0};
            bond<true, true, true>(0, vx.x, vy.x, vz.x, vx.y, vy.y, vz.y, 0.5 * (double)(sa + sb), m2eps, K, sum_e,
                                   sum_cqf, ga, gb);
            {
                const bool in = i + 2u < A.ni;
                if (in || BC == BC_PERIODIC) {
                    double bx, by, bz;
                    int sq;
                    if (in && lane != 63u) {
                        bx = nxx; by = nxy; bz = nxz; sq = gp_lo(nsw);
                    } else {
                        const uint32_t q = in ? l + 2u : l - i;
                        bx = nx[q]; by = ny[q]; bz = nz[q]; sq = s[q];
                    }
                    bond<true, true, false>(0, vx.y, vy.y, vz.y, bx, by, bz, 0.5 * (double)(sb + sq), m2eps, K,
                                            sum_e, sum_cqf, gb, gb);
                }
            }
            {
                const bool in = i >= 1u;
                if (in || BC == BC_PERIODIC) {
                    double bx, by, bz;
                    int sq;
                    if (in && lane != 0u) {
                        bx = pxx; by = pxy; bz = pxz; sq = gp_hi(psw);
                    } else {
                        const uint32_t q = in ? l - 1u : l + A.ni - 1u;
                        bx = nx[q]; by = ny[q]; bz = nz[q]; sq = s[q];
                    }
                    bond<false, false, true>(0, bx, by, bz, vx.x, vy.x, vz.x, 0.5 * (double)(sq + sa), m2eps, K,
                                             sum_e, sum_cqf, ga, ga);
                }
            }
            // y, then z: the pair above (a, b lower ends), then the pair below
            // (a, b upper ends), as in k_energy_grad2
#pragma unroll
            for (int ax = 0; ax < 2; ++ax) {
                const int axis = ax + 1;
                const bool hp = ax == 0 ? hyp : zn(k, 1) < A.nk;
                const bool hm = ax == 0 ? hym : zn(k, -1) < A.nk;
                const GPair &U = ax == 0 ? Yp : (up ? N : P);
                const GPair &D = ax == 0 ? Ym : (up ? P : N);
                if (hp) {
                    const int ta = gp_lo(U.s), tb2 = gp_hi(U.s);
                    bond<true, true, false>(axis, vx.x, vy.x, vz.x, U.x.x, U.y.x, U.z.x, 0.5 * (double)(sa + ta),
                                            m2eps, K, sum_e, sum_cqf, ga, ga);
                    bond<true, true, false>(axis, vx.y, vy.y, vz.y, U.x.y, U.y.y, U.z.y, 0.5 * (double)(sb + tb2),
                                            m2eps, K, sum_e, sum_cqf, gb, gb);
                }
                if (hm) {
                    const int ta = gp_lo(D.s), tb2 = gp_hi(D.s);
                    bond<false, false, true>(axis, D.x.x, D.y.x, D.z.x, vx.x, vy.x, vz.x, 0.5 * (double)(ta + sa),
                                             m2eps, K, sum_e, sum_cqf, ga, ga);
                    bond<false, false, true>(axis, D.x.y, D.y.y, D.z.y, vx.y, vy.y, vz.y, 0.5 * (double)(tb2 + sb),
                                             m2eps, K, sum_e, sum_cqf, gb, gb);
                }
            }
            sum_s += (double)(sa + sb);
            if (A.dEdn) {
                const double pa = fma(ga[2], vz.x, fma(ga[1], vy.x, ga[0] * vx.x));
                const double pb = fma(gb[2], vz.y, fma(gb[1], vy.y, gb[0] * vx.y));
                typedef double d2v __attribute__((ext_vector_type(2)));
                d2v *o = reinterpret_cast<d2v *>(A.dEdn + (size_t)b * 3 * A.N + l);
                __builtin_nontemporal_store(d2v{fma(-pa, vx.x, ga[0]), fma(-pb, vx.y, gb[0])}, o);
                __builtin_nontemporal_store(d2v{fma(-pa, vy.x, ga[1]), fma(-pb, vy.y, gb[1])}, o + A.N / 2);
                __builtin_nontemporal_store(d2v{fma(-pa, vz.x, ga[2]), fma(-pb, vz.y, gb[2])}, o + A.N);
            }
        }
        P = C;
        C = N;
    }
    __shared__ double red[BS / 64][3];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sum_e += __shfl_xor(sum_e, off, 64);
        sum_cqf += __shfl_xor(sum_cqf, off, 64);
        sum_s += __shfl_xor(sum_s, off, 64);
    }
    const uint32_t wid = tid >> 6;
    if (lane == 0) { red[wid][0] = sum_e; red[wid][1] = sum_cqf; red[wid][2] = sum_s; }
    __syncthreads();
    if (tid < 3) {
        double v = 0.0;
        for (int w = 0; w < BS / 64; ++w) v += red[w][tid];
        A.part[((size_t)b * A.nblk + x) * 4 + tid] = v;
    }
}

// One block per lattice: E = eps S_e - K S_cqf - g S_s and the parameter
// derivatives dE/deps = S_e, dE/dK = -S_cqf, dE/dg = -S_s.
__global__ __launch_bounds__(256) void k_energy_final(GradArgs A, double *E, double *dK, double *dg, double *deps)
{
    const uint32_t b = blockIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    for (uint32_t q = threadIdx.x; q < A.nblk; q += 256) {
        const double *p = A.part + ((size_t)b * A.nblk + q) * 4;
        v[0] += p[0]; v[1] += p[1]; v[2] += p[2];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] += __shfl_xor(v[c], off, 64);
    __shared__ double red[4][3];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (lane == 0) { red[wid][0] = v[0]; red[wid][1] = v[1]; red[wid][2] = v[2]; }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double t[3];
    for (int c = 0; c < 3; ++c) t[c] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
    const double K = A.K[b], g = A.g[b], eps = A.eps[b];
    if (E) E[b] = eps * t[0] - K * t[1] - g * t[2];
    if (deps) deps[b] = t[0];
    if (dK) dK[b] = -t[1];
    if (dg) dg[b] = -t[2];
}

}  // namespace chll

using namespace chll;

// Workspace of chll_energy_grad_noalloc: CHLL_ENERGY_SCRATCH_BYTES per GPU in
// the code object's own device memory, held from the library's first use on
// that GPU until the process ends.  Host code must not pass `scratch` itself
// to a kernel (in host code it names a host-side stand-in): it uses the
// device address that hipGetSymbolAddress returns, looked up once per device
// with that device current (scratch_address).
static __device__ __align__(256) double scratch[CHLL_ENERGY_SCRATCH_BYTES / sizeof(double)];

namespace {
thread_local std::string g_gerr;
int gfail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_gerr = buf;
    return -1;
}
}  // namespace

extern "C" {

const char *chll_energy_last_error(void) { return g_gerr.c_str(); }

size_t chll_energy_workspace_bytes(int32_t batch, int32_t ni, int32_t nj, int32_t nk)
{
    const uint64_t N = (uint64_t)ni * nj * nk;
    const uint64_t nblk = (N + kGradBlock - 1) / kGradBlock;
    return (size_t)batch * nblk * 4 * sizeof(double);
}

}  // extern "C"

namespace {

int check_grad_args(const char *who, const double *n, const int8_t *s, int32_t batch, int32_t ni, int32_t nj,
                    int32_t nk, int32_t bc, const double *K, const double *g, const double *eps)
{
    if (!n || !s || !K || !g || !eps) return gfail("%s: null argument", who);
    if (batch < 1 || batch > 65535 || ni < 1 || nj < 1 || nk < 1) return gfail("%s: bad shape", who);
    if (bc != BC_OPEN && bc != BC_PERIODIC) return gfail("%s: unknown boundary condition %d", who, bc);
    const uint64_t N = (uint64_t)ni * nj * nk;
    if (N >= (1ull << 32) / 3) return gfail("%s: lattice too large", who);
    return 0;
}

// Both entry points: the gather kernel writes every block's partial sums into
// part (batch x nblk x 4 doubles, each written once before it is read), the
// final kernel reduces them in a fixed order.
int grad_launch(const double *n, const int8_t *s, int32_t batch, int32_t ni, int32_t nj, int32_t nk, int32_t bc,
                const double *K, const double *g, const double *eps, double *E, double *dEdn, double *dEdK,
                double *dEdg, double *dEdeps, double *part, void *stream)
{
    const uint64_t N = (uint64_t)ni * nj * nk;
    GradArgs A;
    A.n = n; A.s = s; A.K = K; A.g = g; A.eps = eps; A.dEdn = dEdn; A.part = part;
    A.ni = ni; A.nj = nj; A.nk = nk; A.N = (uint32_t)N;
    A.fd_ni = make_fastdiv(ni); A.fd_nj = make_fastdiv(nj); A.bc = bc;
    // two sites per thread when every row splits into aligned pairs
    const bool pairs = ni % 2 == 0 && (uintptr_t)n % 16 == 0 && (uintptr_t)dEdn % 16 == 0 && (uintptr_t)s % 2 == 0;
    const uint64_t per_thread = pairs ? 2 : 1;
    A.nblk = (uint32_t)((N / per_thread + kGradBlock - 1) / kGradBlock);
    // XCD row bands when a plane is 8 bands of whole blocks (then every
    // lattice's block count is a multiple of 8 and block (x, b) runs on XCD x % 8)
    const uint64_t plane = (uint64_t)ni * nj / per_thread;
    A.band = plane % (8u * kGradBlock) == 0 ? (uint32_t)(plane / kGradBlock) : 0u;
    hipStream_t st = (hipStream_t)stream;
    // marching slabs of ks planes (k_energy_grad_march) when the blocks fit the
    // workspace that chll_energy_workspace_bytes sizes (one partial per block
    // of one-site threads); blocks of 1024 threads (taller tiles) ran slower
    constexpr uint32_t kMarchBlock = kGradBlock;
    const uint32_t tiles = (uint32_t)((plane + kMarchBlock - 1) / kMarchBlock);
    // slabs of 8 planes (16 ran the same at 8 x 128^3; 32 slower, with more
    // re-reads)
    const uint32_t ks = nk < 8 ? (uint32_t)nk : 8u;
    const uint32_t nslab = ((uint32_t)nk + ks - 1) / ks;
    if (pairs && (uint64_t)tiles * nslab <= (N + kGradBlock - 1) / kGradBlock) {
        A.nblk = tiles * nslab;
        A.band = 0;
        // a batch of whole multiples of 8 lattices: one lattice per XCD at a time
        const int whole = batch % 8 == 0;
        const dim3 grid = whole ? dim3(A.nblk * (uint32_t)batch) : dim3(A.nblk, batch), block(kMarchBlock);
        if (bc == BC_OPEN)
            hipLaunchKernelGGL((k_energy_grad_march<BC_OPEN, kMarchBlock>), grid, block, 0, st, A, tiles, nslab, ks,
                               whole);
        else
            hipLaunchKernelGGL((k_energy_grad_march<BC_PERIODIC, kMarchBlock>), grid, block, 0, st, A, tiles, nslab,
                               ks, whole);
        hipLaunchKernelGGL(k_energy_final, dim3(batch), dim3(256), 0, st, A, E, dEdK, dEdg, dEdeps);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return gfail("chll_energy_grad launch: %s", hipGetErrorString(e));
        return 0;
    }
    dim3 grid(A.nblk, batch);
    if (pairs) {
        if (bc == BC_OPEN) hipLaunchKernelGGL(k_energy_grad2<BC_OPEN>, grid, dim3(kGradBlock), 0, st, A);
        else hipLaunchKernelGGL(k_energy_grad2<BC_PERIODIC>, grid, dim3(kGradBlock), 0, st, A);
    } else {
        if (bc == BC_OPEN) hipLaunchKernelGGL(k_energy_grad<BC_OPEN>, grid, dim3(kGradBlock), 0, st, A);
        else hipLaunchKernelGGL(k_energy_grad<BC_PERIODIC>, grid, dim3(kGradBlock), 0, st, A);
    }
    hipLaunchKernelGGL(k_energy_final, dim3(batch), dim3(256), 0, st, A, E, dEdK, dEdg, dEdeps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return gfail("chll_energy_grad launch: %s", hipGetErrorString(e));
    return 0;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

int scratch_address(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return gfail("chll_energy_grad_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return gfail("chll_energy_grad_noalloc: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return gfail("chll_energy_grad_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_dev[dev] = (double *)p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

}  // namespace

extern "C" {

int chll_energy_grad(const double *n, const int8_t *s, int32_t batch, int32_t ni, int32_t nj, int32_t nk, int32_t bc,
                     const double *K, const double *g, const double *eps, double *E, double *dEdn, double *dEdK,
                     double *dEdg, double *dEdeps, void *workspace, size_t workspace_bytes, void *stream)
{
    g_gerr.clear();
    if (check_grad_args("chll_energy_grad", n, s, batch, ni, nj, nk, bc, K, g, eps)) return -1;
    if (!workspace) return gfail("chll_energy_grad: null argument");
    if (workspace_bytes < chll_energy_workspace_bytes(batch, ni, nj, nk)) return gfail("chll_energy_grad: workspace too small");
    return grad_launch(n, s, batch, ni, nj, nk, bc, K, g, eps, E, dEdn, dEdK, dEdg, dEdeps, (double *)workspace, stream);
}

int chll_energy_grad_noalloc(const double *n, const int8_t *s, int32_t batch, int32_t ni, int32_t nj, int32_t nk,
                             int32_t bc, const double *K, const double *g, const double *eps, double *E, double *dEdn,
                             double *dEdK, double *dEdg, double *dEdeps, void *stream)
{
    g_gerr.clear();
    if (check_grad_args("chll_energy_grad_noalloc", n, s, batch, ni, nj, nk, bc, K, g, eps)) return -1;
    const size_t need = chll_energy_workspace_bytes(batch, ni, nj, nk);
    if (need > CHLL_ENERGY_SCRATCH_BYTES)
        return gfail("chll_energy_grad_noalloc: batch %d of %dx%dx%d needs %zu bytes of scratch, the static buffer "
                     "holds %zu (use chll_energy_grad with a workspace)",
                     batch, ni, nj, nk, need, (size_t)CHLL_ENERGY_SCRATCH_BYTES);
    double *part = nullptr;
    if (scratch_address(&part)) return -1;
    return grad_launch(n, s, batch, ni, nj, nk, bc, K, g, eps, E, dEdn, dEdK, dEdg, dEdeps, part, stream);
}

}  // extern "C"
