// chll_kernels.hip -- HIP kernels (gfx950) of the chiral Lebwohl-Lasher
// Metropolis sweep and their host-side launchers.
//
// One sweep = four dependent half-sweeps over the colour-split lattice
// (director colour 0, director colour 1, chirality colour 0, chirality colour
// 1; chll/chll_mod.f90:70-81), then a one-block finalize kernel that sums the
// per-block accept counters, adapts the cone step d on the device
// (chll/chll_mod.f90:83-90) and, on measurement sweeps, closes the
// observable reductions into one row (chll/chll_mod.f90:91-105).  Nothing in a
// run of sweeps waits for the host.
#include "chll_kernels.h"

namespace chll {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Layout transforms: Fortran-ordered host fields <-> colour-split SoA.
__global__ __launch_bounds__(kBlock) void k_scatter_f64(Geom G, const double *__restrict__ src,
                                                        double *__restrict__ dst)
{
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= G.nsub) return;
    const uint32_t row = fdiv(m, G.fd_half);
    const uint32_t k = fdiv(row, G.fd_nj);
    const uint32_t j = row - k * G.nj;
    const uint32_t c0 = (j + k) & 1u;  // colour of the site with even i (l = 2m)
    const double a = src[2ull * m], b = src[2ull * m + 1];
    dst[c0 * G.nsub + m] = a;
    dst[(c0 ^ 1u) * G.nsub + m] = b;
}

__global__ __launch_bounds__(kBlock) void k_gather_f64(Geom G, const double *__restrict__ src,
                                                       double *__restrict__ dst)
{
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= G.nsub) return;
    const uint32_t row = fdiv(m, G.fd_half);
    const uint32_t k = fdiv(row, G.fd_nj);
    const uint32_t j = row - k * G.nj;
    const uint32_t c0 = (j + k) & 1u;
    dst[2ull * m] = src[c0 * G.nsub + m];
    dst[2ull * m + 1] = src[(c0 ^ 1u) * G.nsub + m];
}

__global__ __launch_bounds__(kBlock) void k_scatter_spin(Geom G, const int32_t *__restrict__ src,
                                                         int8_t *__restrict__ dst, int *__restrict__ bad)
{
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= G.nsub) return;
    const uint32_t row = fdiv(m, G.fd_half);
    const uint32_t k = fdiv(row, G.fd_nj);
    const uint32_t j = row - k * G.nj;
    const uint32_t c0 = (j + k) & 1u;
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
    const uint32_t row = fdiv(m, G.fd_half);
    const uint32_t k = fdiv(row, G.fd_nj);
    const uint32_t j = row - k * G.nj;
    const uint32_t c0 = (j + k) & 1u;
    dst[2ull * m] = src[c0 * G.nsub + m];
    dst[2ull * m + 1] = src[(c0 ^ 1u) * G.nsub + m];
}

// ---------------------------------------------------------------------------
// Block-wide sum of one 0/1 flag per thread, added to slot blk % kSlots.
__device__ inline void block_count(bool flag, uint32_t *slots, uint32_t blk)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const uint64_t bal = __ballot(flag);
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (lane == 0) wsum[wid] = (uint32_t)__popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t += wsum[w];
        if (t) atomicAdd(&slots[(blk & (kSlots - 1)) * kSlotStride], t);
    }
}

__device__ inline void load_nbrs(const Fields &F, uint32_t oth, const uint32_t nb[6], Nbr &N)
{
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const uint32_t q = oth + nb[b];
        N.x[b] = F.nx[q];
        N.y[b] = F.ny[q];
        N.z[b] = F.nz[q];
        N.s[b] = F.s[q];
    }
}

// Director half-sweep over one colour (chll/chll_mod.f90:117-202).
template <int ARITH, int RNG, int BC>
__global__ __launch_bounds__(kBlock) void k_dir(Geom G, Fields F, int colour, PhaseArgs A,
                                                uint32_t *__restrict__ counts)
{
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t m = blk * kBlock + threadIdx.x;
    bool acc = false;
    if (m < G.nsub) {
        const Site st = decode(G, m, colour);
        uint32_t nb[6];
        Nbr N;
        N.mask = neighbours<BC>(G, m, st, nb);
        const uint32_t own = colour ? G.nsub + m : m;
        const uint32_t oth = colour ? 0u : G.nsub;
        load_nbrs(F, oth, nb, N);
        const double nx = F.nx[own], ny = F.ny[own], nz = F.nz[own];
        const int sg = F.s[own];
        const double d = *A.d;
        double u_phi, u_acc;
        if (RNG == RNG_PHILOX) {
            const U4 w = site_block(A.seed, A.run_index, A.replica, m, A.istep, colour, 0);
            u_phi = u53(w.y, w.x);
            u_acc = u53(w.w, w.z);
        } else {
            u_phi = A.u_phi[m];
            u_acc = A.u_acc[m];
        }
        double tx, ty, tz;
        if (ARITH == ARITH_COMPAT) {
            const double eold = energy_compat(N, nx, ny, nz, sg, A.g, A.KK);
            trial_compat(nx, ny, nz, d, u_phi, tx, ty, tz);
            const double enew = energy_compat(N, tx, ty, tz, sg, A.g, A.KK);
            acc = (enew < eold) || (u_acc <= exp(div_by_recip(-(enew - eold), A.kbt, A.inv_kbt)));
        } else {
            double c2o, c2n;
            const double ao = cqf_sum(N, sg, nx, ny, nz, c2o);
            trial_fp64(nx, ny, nz, d, u_phi, tx, ty, tz);
            const double an = cqf_sum(N, sg, tx, ty, tz, c2n);
            const double dE = (c2o - c2n) - A.KK * (an - ao);
            acc = (dE < 0.0) || (u_acc <= exp(-dE * A.inv_kbt));
        }
        if (acc) {
            F.nx[own] = tx;
            F.ny[own] = ty;
            F.nz[own] = tz;
        }
    }
    block_count(acc, counts, blk);
}

// Chirality half-sweep over one colour (chll/chll_mod.f90:204-245).
template <int ARITH, int RNG, int BC>
__global__ __launch_bounds__(kBlock) void k_spin(Geom G, Fields F, int colour, PhaseArgs A,
                                                 uint32_t *__restrict__ counts)
{
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t m = blk * kBlock + threadIdx.x;
    bool acc = false;
    if (m < G.nsub) {
        const Site st = decode(G, m, colour);
        uint32_t nb[6];
        Nbr N;
        N.mask = neighbours<BC>(G, m, st, nb);
        const uint32_t own = colour ? G.nsub + m : m;
        const uint32_t oth = colour ? 0u : G.nsub;
        load_nbrs(F, oth, nb, N);
        const double nx = F.nx[own], ny = F.ny[own], nz = F.nz[own];
        const int sg = F.s[own];
        double u_acc;
        if (RNG == RNG_PHILOX) {
            const U4 w = site_block(A.seed, A.run_index, A.replica, m, A.istep, colour, 1);
            u_acc = u53(w.y, w.x);
        } else {
            u_acc = A.u_acc[m];
        }
        if (ARITH == ARITH_COMPAT) {
            const double eold = energy_compat(N, nx, ny, nz, sg, A.g, A.KK);
            const double enew = energy_compat(N, nx, ny, nz, -sg, A.g, A.KK);
            acc = (enew < eold) || (u_acc <= exp(div_by_recip(-(enew - eold), A.kbt, A.inv_kbt)));
        } else {
            double c2;
            const double a = cq_sum(N, nx, ny, nz, c2);
            const double dE = (double)sg * fma(A.KK, a, 2.0 * A.g);
            acc = (dE < 0.0) || (u_acc <= exp(-dE * A.inv_kbt));
        }
        if (acc) F.s[own] = (int8_t)(-sg);
    }
    block_count(acc, counts, blk);
}

// ---------------------------------------------------------------------------
// Observables.  Per site: the reference's reported energy term (periodic i
// and j, open k, +y bond x-component from the diagonal site (i+1, j+1, k),
// chll/chll_mod.f90:427-452), the Hamiltonian's own +x/+y/+z bonds under the
// sweep's boundary condition, the six Q-tensor sums n_a n_b and the spin.
__device__ inline uint32_t pidx(const Geom &G, uint32_t i, uint32_t j, uint32_t k)
{
    return (i >> 1) + G.half * (j + G.nj * k);
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

template <int BC>
__global__ __launch_bounds__(kBlock) void k_measure(Geom G, Fields F, double g, double KK,
                                                    double *__restrict__ part, long long *__restrict__ spart)
{
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = 0.0;
    long long ssum = 0;
    for (uint32_t m = blockIdx.x * kBlock + threadIdx.x; m < G.nsub; m += gridDim.x * kBlock) {
        const uint32_t row = fdiv(m, G.fd_half);
        const uint32_t h = m - row * G.half;
        const uint32_t k = fdiv(row, G.fd_nj);
        const uint32_t j = row - k * G.nj;
        const uint32_t jp = (j + 1 == G.nj) ? 0u : j + 1;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const uint32_t i = 2u * h + (uint32_t)b;
            const uint32_t ip = (i + 1 == G.ni) ? 0u : i + 1;
            const uint32_t c = colour_base(G, i, j, k) + m;
            const double ax = F.nx[c], ay = F.ny[c], az = F.nz[c];
            const int sa = F.s[c];
            // +x neighbour (periodic wrap in i)
            const uint32_t qx = colour_base(G, ip, j, k) + pidx(G, ip, j, k);
            const double bx0 = F.nx[qx], by0 = F.ny[qx], bz0 = F.nz[qx];
            const int sx = F.s[qx];
            // +y neighbour (periodic wrap in j) and the diagonal site's x component
            const uint32_t qy = colour_base(G, i, jp, k) + pidx(G, i, jp, k);
            const double bx1 = F.nx[qy], by1 = F.ny[qy], bz1 = F.nz[qy];
            const int sy = F.s[qy];
            const uint32_t qd = colour_base(G, ip, jp, k) + pidx(G, ip, jp, k);
            const double dgx = F.nx[qd];
            const double tx = bond_term(ax, ay, az, bx0, by0, bz0, 0, sa, sx, KK);
            // reference's +y bond: dot uses nx of the diagonal site
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
            v[0] += ((tx + ety) + tz) - gs;
            const bool hx = (BC == BC_PERIODIC) || (i + 1 < G.ni);
            const bool hy = (BC == BC_PERIODIC) || (j + 1 < G.nj);
            v[1] += ((hx ? tx : 0.0) + (hy ? ty : 0.0) + tzp) - gs;
            v[2] += ax * ax;
            v[3] += ay * ay;
            v[4] += az * az;
            v[5] += ax * ay;
            v[6] += ax * az;
            v[7] += ay * az;
            ssum += sa;
        }
    }
    // deterministic block reduction: wave butterflies, then a fixed-order LDS pass
    __shared__ double wv[kBlock / 64][8];
    __shared__ long long ws[kBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] += __shfl_xor(v[q], off, 64);
        ssum += __shfl_xor(ssum, off, 64);
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) wv[wid][q] = v[q];
        ws[wid] = ssum;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += wv[w][threadIdx.x];
        part[(size_t)blockIdx.x * 8 + threadIdx.x] = t;
    }
    if (threadIdx.x == 8) {
        long long t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += ws[w];
        spart[blockIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// Sweep epilogue (one block): accept/flip totals, step adaptation, measurement
// row.  Row layout (kRowCols doubles): istep, E (reference etot), eta,
// paccept, faccept, d, E_true/N, eta_exact, Qxx, Qyy, Qzz, Qxy, Qxz, Qyz
// (means of n_a n_b), naccept, nflip.
constexpr int kFinBlock = 256;

__global__ __launch_bounds__(kFinBlock) void k_finalize(FinArgs A)
{
    __shared__ double red[kFinBlock / 64][10];
    const uint32_t t = threadIdx.x;
    unsigned long long ca = 0, cf = 0;
    for (uint32_t q = t; q < 4 * kSlots; q += kFinBlock) {
        uint32_t *slot = A.counts + (q / kSlots) * (kSlots * kSlotStride) + (q % kSlots) * kSlotStride;
        const uint32_t c = *slot;
        *slot = 0u;
        if (q < 2 * kSlots) ca += c;
        else cf += c;
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
    double cad = (double)ca, cfd = (double)cf, ssd = (double)ss;  // exact: < 2^53
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] += __shfl_xor(v[q], off, 64);
        cad += __shfl_xor(cad, off, 64);
        cfd += __shfl_xor(cfd, off, 64);
        ssd += __shfl_xor(ssd, off, 64);
    }
    const uint32_t lane = t & 63u, wid = t >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) red[wid][q] = v[q];
        red[wid][8] = cad;
        red[wid][9] = cfd;
    }
    __shared__ double sred[kFinBlock / 64];
    if (lane == 0) sred[wid] = ssd;
    __syncthreads();
    if (t != 0) return;
    double tot[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) {
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
    if (A.counts_out) {
        A.counts_out[2 * (size_t)A.sweep_index] = (long long)tot[8];
        A.counts_out[2 * (size_t)A.sweep_index + 1] = (long long)tot[9];
    }
    if (A.measure) {
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
// Launchers.
static inline dim3 grid1(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

uint32_t phase_blocks(const Geom &G) { return (G.nsub + kBlock - 1) / kBlock; }

uint32_t measure_blocks(const Geom &G)
{
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

template <int ARITH, int RNG, int BC>
static void launch_phase_t(int kind, const Geom &G, const Fields &F, int colour, const PhaseArgs &A,
                           uint32_t *counts, hipStream_t s)
{
    if (kind == 0)
        hipLaunchKernelGGL((k_dir<ARITH, RNG, BC>), grid1(G.nsub), dim3(kBlock), 0, s, G, F, colour, A, counts);
    else
        hipLaunchKernelGGL((k_spin<ARITH, RNG, BC>), grid1(G.nsub), dim3(kBlock), 0, s, G, F, colour, A, counts);
}

void launch_phase(int kind, int arith, int rng, int bc, const Geom &G, const Fields &F, int colour,
                  const PhaseArgs &A, uint32_t *counts, hipStream_t s)
{
    const bool philox = (rng == RNG_PHILOX);
    if (arith == ARITH_COMPAT) {
        if (philox) {
            if (bc == BC_OPEN) launch_phase_t<ARITH_COMPAT, RNG_PHILOX, BC_OPEN>(kind, G, F, colour, A, counts, s);
            else launch_phase_t<ARITH_COMPAT, RNG_PHILOX, BC_PERIODIC>(kind, G, F, colour, A, counts, s);
        } else {
            if (bc == BC_OPEN) launch_phase_t<ARITH_COMPAT, RNG_STREAM, BC_OPEN>(kind, G, F, colour, A, counts, s);
            else launch_phase_t<ARITH_COMPAT, RNG_STREAM, BC_PERIODIC>(kind, G, F, colour, A, counts, s);
        }
    } else {
        if (philox) {
            if (bc == BC_OPEN) launch_phase_t<ARITH_FP64, RNG_PHILOX, BC_OPEN>(kind, G, F, colour, A, counts, s);
            else launch_phase_t<ARITH_FP64, RNG_PHILOX, BC_PERIODIC>(kind, G, F, colour, A, counts, s);
        } else {
            if (bc == BC_OPEN) launch_phase_t<ARITH_FP64, RNG_STREAM, BC_OPEN>(kind, G, F, colour, A, counts, s);
            else launch_phase_t<ARITH_FP64, RNG_STREAM, BC_PERIODIC>(kind, G, F, colour, A, counts, s);
        }
    }
}

void launch_measure(int bc, const Geom &G, const Fields &F, double g, double KK, double *part, long long *spart,
                    hipStream_t s)
{
    if (bc == BC_OPEN)
        hipLaunchKernelGGL(k_measure<BC_OPEN>, dim3(measure_blocks(G)), dim3(kBlock), 0, s, G, F, g, KK, part, spart);
    else
        hipLaunchKernelGGL(k_measure<BC_PERIODIC>, dim3(measure_blocks(G)), dim3(kBlock), 0, s, G, F, g, KK, part, spart);
}

void launch_finalize(const FinArgs &A, hipStream_t s)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, A);
}

void launch_stream_fill(double *out, uint64_t seed, uint64_t base, uint32_t nsub, hipStream_t s)
{
    dim3 grid((nsub + kBlock - 1) / kBlock, 6);
    hipLaunchKernelGGL(k_stream_fill, grid, dim3(kBlock), 0, s, out, seed, base, nsub);
}

}  // namespace chll
