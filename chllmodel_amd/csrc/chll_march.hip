// chll_march.hip -- half-sweep kernels that march along k through LDS.
//
// A block owns a tile of TH packed sites x TJ rows of the (h, j) plane and a
// slab of planes [k0, k1).  For every plane k of the slab it updates the
// active colour's sites of the tile; their six neighbours are all of the
// other colour: at plane k the packed sites (h+p-1, j), (h+p, j), (h, j-1),
// (h, j+1) (p = i & 1), and the site (h, j) of planes k-1 and k+1.  The block
// keeps the other colour's plane tiles for k-1, k, k+1, each with a one-site
// halo, in an LDS ring, and loads plane k+2 into registers while it computes
// plane k, so every other-colour value is fetched from memory once per slab
// pass (plus halo) instead of once per neighbouring site.  The active site's
// own director and spin stream through registers one plane ahead.
//
// The trial arithmetic (chll_site.h dir_compute / spin_compute) and random
// counters are those of the one-site-per-thread kernels, so both produce
// identical lattices.  Only the active colour is written, and only the other
// colour is read through LDS, so a half-sweep has no read-after-write hazard.
#include "chll_kernels.h"
#include "chll_site.h"

namespace chll {

// Neighbour accessor reading the LDS ring at every use: flat index of each of
// the six neighbours into the [3][NS] tile arrays.
struct LdsNbr {
    const double *lx, *ly, *lz;
    const int8_t *ls;
    int idx[6];
    uint32_t mask;
    __device__ void get(int b, double &mx, double &my, double &mz) const
    {
        mx = lx[idx[b]];
        my = ly[idx[b]];
        mz = lz[idx[b]];
    }
    __device__ int spin(int b) const { return ls[idx[b]]; }
};

template <int KIND, int ARITH, int RNG, int BC, int TH, int TJ>
__global__ __launch_bounds__(TH * TJ) void k_march(Geom G, Fields F, int colour, PhaseArgs A,
                                                   uint32_t *__restrict__ counts, uint32_t tiles_x, uint32_t slab)
{
    constexpr int NT = TH * TJ;
    constexpr int W = TH + 2, H = TJ + 2, NS = W * H;
    constexpr int PER = (NS + NT - 1) / NT;
    __shared__ double lx[3][NS], ly[3][NS], lz[3][NS];
    __shared__ int8_t ls[3][NS];

    const uint32_t tid = threadIdx.x;
    const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int h0 = (int)(tx * TH), j0 = (int)(ty * TJ);
    const uint32_t k0 = blockIdx.y * slab;
    const uint32_t k1 = (k0 + slab < G.nk) ? k0 + slab : G.nk;
    const uint32_t oth = colour ? 0u : G.nsub;
    const uint32_t ownb = colour ? G.nsub : 0u;
    const uint32_t hl = tid % TH, jl = tid / TH;
    const uint32_t h = (uint32_t)h0 + hl, j = (uint32_t)j0 + jl;
    const bool in = (h < G.half) && (j < G.nj);
    const int half = (int)G.half, nj = (int)G.nj, nk = (int)G.nk;

    double px[PER], py[PER], pz[PER];
    int ps[PER];
    // Other-colour halo tile of plane kk into registers (zero outside an open lattice).
    auto fetch = [&](int kk) {
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int q = (int)tid + r * NT;
            double vx = 0.0, vy = 0.0, vz = 0.0;
            int vs = 0;
            if (q < NS) {
                int hh = h0 - 1 + q % W, jj = j0 - 1 + q / W, kq = kk;
                bool ok;
                if (BC == BC_PERIODIC) {
                    if (hh < 0) hh += half; else if (hh >= half) hh -= half;
                    if (hh >= half) hh = 0;
                    if (jj < 0) jj += nj; else if (jj >= nj) jj -= nj;
                    if (jj >= nj) jj = 0;
                    if (kq < 0) kq += nk; else if (kq >= nk) kq -= nk;
                    ok = true;
                } else {
                    ok = hh >= 0 && hh < half && jj >= 0 && jj < nj && kq >= 0 && kq < nk;
                }
                if (ok) {
                    const uint32_t idx = oth + (uint32_t)hh + G.half * ((uint32_t)jj + G.nj * (uint32_t)kq);
                    vx = F.nx[idx];
                    vy = F.ny[idx];
                    vz = F.nz[idx];
                    vs = F.s[idx];
                }
            }
            px[r] = vx; py[r] = vy; pz[r] = vz; ps[r] = vs;
        }
    };
    auto store = [&](int slot) {
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int q = (int)tid + r * NT;
            if (q < NS) {
                lx[slot][q] = px[r];
                ly[slot][q] = py[r];
                lz[slot][q] = pz[r];
                ls[slot][q] = (int8_t)ps[r];
            }
        }
    };

    // prime the ring with planes k0-1, k0, k0+1
    int sm = 0, sc = 1, sp = 2;  // slots of planes k-1, k, k+1
    fetch((int)k0 - 1);
    store(sm);
    fetch((int)k0);
    store(sc);
    fetch((int)k0 + 1);
    store(sp);
    // own site of plane k0
    double ox = 0.0, oy = 0.0, oz = 0.0;
    int os = 0;
    auto fetch_own = [&](uint32_t kk) {
        if (in && kk < k1) {
            const uint32_t idx = ownb + h + G.half * (j + G.nj * kk);
            ox = F.nx[idx];
            oy = F.ny[idx];
            oz = F.nz[idx];
            os = F.s[idx];
        }
    };
    fetch_own(k0);
    __syncthreads();

    const double d = (KIND == 0) ? *A.d : 0.0;
    const int cx = (int)hl + 1, cy = (int)jl + 1;
    uint32_t cnt = 0;
    for (uint32_t k = k0; k < k1; ++k) {
        const double snx = ox, sny = oy, snz = oz;
        const int ssg = os;
        const bool more = k + 1 < k1;
        fetch_own(k + 1);
        if (more) fetch((int)k + 2);
        if (in) {
            const uint32_t p = (j + k + (uint32_t)colour) & 1u;
            const int c0 = cy * W + cx, xp = c0 + (int)p;
            LdsNbr N;
            N.lx = &lx[0][0]; N.ly = &ly[0][0]; N.lz = &lz[0][0]; N.ls = &ls[0][0];
            N.idx[0] = sc * NS + xp;
            N.idx[1] = sc * NS + xp - 1;
            N.idx[2] = sc * NS + c0 + W;
            N.idx[3] = sc * NS + c0 - W;
            N.idx[4] = sp * NS + c0;
            N.idx[5] = sm * NS + c0;
            const uint32_t i = 2u * h + p;
            if (BC == BC_PERIODIC) {
                N.mask = 63u;
            } else {
                N.mask = (i + 1 < G.ni ? 1u : 0u) | (i >= 1 ? 2u : 0u) | (j + 1 < G.nj ? 4u : 0u) |
                         (j >= 1 ? 8u : 0u) | (k + 1 < G.nk ? 16u : 0u) | (k >= 1 ? 32u : 0u);
            }
            const uint32_t m = h + G.half * (j + G.nj * k);
            const uint32_t own = ownb + m;
            if (KIND == 0) {
                double t[3];
                if (dir_compute<ARITH, RNG>(A, colour, m, N, snx, sny, snz, ssg, d, t)) {
                    F.nx[own] = t[0];
                    F.ny[own] = t[1];
                    F.nz[own] = t[2];
                    ++cnt;
                }
            } else {
                if (spin_compute<ARITH, RNG>(A, colour, m, N, snx, sny, snz, ssg)) {
                    F.s[own] = (int8_t)(-ssg);
                    ++cnt;
                }
            }
        }
        __syncthreads();  // everyone is done with plane k-1's slot
        if (more) store(sm);
        const int t = sm;
        sm = sc;
        sc = sp;
        sp = t;
        __syncthreads();
    }
    // block total of accepted trials
    __shared__ uint32_t wsum[NT / 64];
    uint32_t v = cnt;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((tid & 63u) == 0) wsum[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) s += wsum[w];
        if (s) atomicAdd(&counts[((blockIdx.x + blockIdx.y * gridDim.x) & (kSlots - 1)) * kSlotStride], s);
    }
}


// Single-plane tile variant: a block updates one TH x TJ tile of one plane.
// The other colour's in-plane neighbours come from an LDS copy of the tile
// and its one-site halo, loaded once by the block; the two z-neighbours and
// the site itself are read directly.  No state is carried between planes, so
// the kernel keeps the one-site kernel's register budget.
template <int KIND, int ARITH, int RNG, int BC, int TH, int TJ>
__global__ __launch_bounds__(TH * TJ) void k_tile(Geom G, Fields F, int colour, PhaseArgs A,
                                                  uint32_t *__restrict__ counts, uint32_t tiles_x)
{
    constexpr int NT = TH * TJ;
    constexpr int W = TH + 2, H = TJ + 2, NS = W * H;
    constexpr int PER = (NS + NT - 1) / NT;
    __shared__ double lx[NS], ly[NS], lz[NS];
    __shared__ int8_t ls[NS];

    const uint32_t tid = threadIdx.x;
    const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int h0 = (int)(tx * TH), j0 = (int)(ty * TJ);
    const uint32_t k = blockIdx.y;
    const uint32_t oth = colour ? 0u : G.nsub;
    const uint32_t ownb = colour ? G.nsub : 0u;
    const uint32_t hl = tid % TH, jl = tid / TH;
    const uint32_t h = (uint32_t)h0 + hl, j = (uint32_t)j0 + jl;
    const bool in = (h < G.half) && (j < G.nj);
    const int half = (int)G.half, nj = (int)G.nj;
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int q = (int)tid + r * NT;
        if (q < NS) {
            int hh = h0 - 1 + q % W, jj = j0 - 1 + q / W;
            bool ok;
            if (BC == BC_PERIODIC) {
                if (hh < 0) hh += half; else if (hh >= half) hh -= half;
                if (hh >= half) hh = 0;
                if (jj < 0) jj += nj; else if (jj >= nj) jj -= nj;
                if (jj >= nj) jj = 0;
                ok = true;
            } else {
                ok = hh >= 0 && hh < half && jj >= 0 && jj < nj;
            }
            const uint32_t idx = ok ? oth + (uint32_t)hh + G.half * ((uint32_t)jj + G.nj * k) : G.zero;
            lx[q] = F.nx[idx];
            ly[q] = F.ny[idx];
            lz[q] = F.nz[idx];
            ls[q] = F.s[idx];
        }
    }
    __syncthreads();
    bool acc = false;
    if (in) {
        const uint32_t p = (j + k + (uint32_t)colour) & 1u;
        const int c0 = ((int)jl + 1) * W + (int)hl + 1, xp = c0 + (int)p;
        const int q[4] = {xp, xp - 1, c0 + W, c0 - W};
        Nbr N;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            N.x[b] = lx[q[b]];
            N.y[b] = ly[q[b]];
            N.z[b] = lz[q[b]];
            N.s[b] = ls[q[b]];
        }
        const uint32_t m = h + G.half * (j + G.nj * k);
        const uint32_t i = 2u * h + p;
        // z-neighbours: the other colour at (h, j) of planes k+1 and k-1
        uint32_t zp, zm;
        if (BC == BC_PERIODIC) {
            zp = oth + (k + 1 < G.nk ? m + G.plane : m - (G.nk - 1) * G.plane);
            zm = oth + (k >= 1 ? m - G.plane : m + (G.nk - 1) * G.plane);
            N.mask = 63u;
        } else {
            zp = k + 1 < G.nk ? oth + m + G.plane : G.zero;
            zm = k >= 1 ? oth + m - G.plane : G.zero;
            N.mask = (i + 1 < G.ni ? 1u : 0u) | (i >= 1 ? 2u : 0u) | (j + 1 < G.nj ? 4u : 0u) | (j >= 1 ? 8u : 0u) |
                     (k + 1 < G.nk ? 16u : 0u) | (k >= 1 ? 32u : 0u);
        }
        N.x[4] = F.nx[zp]; N.y[4] = F.ny[zp]; N.z[4] = F.nz[zp]; N.s[4] = F.s[zp];
        N.x[5] = F.nx[zm]; N.y[5] = F.ny[zm]; N.z[5] = F.nz[zm]; N.s[5] = F.s[zm];
        const uint32_t own = ownb + m;
        const double nx = F.nx[own], ny = F.ny[own], nz = F.nz[own];
        const int sg = F.s[own];
        if (KIND == 0) {
            double t[3];
            acc = dir_compute<ARITH, RNG>(A, colour, m, N, nx, ny, nz, sg, *A.d, t);
            if (acc) {
                F.nx[own] = t[0];
                F.ny[own] = t[1];
                F.nz[own] = t[2];
            }
        } else {
            acc = spin_compute<ARITH, RNG>(A, colour, m, N, nx, ny, nz, sg);
            if (acc) F.s[own] = (int8_t)(-sg);
        }
    }
    block_count<NT>(acc, counts, blockIdx.x + blockIdx.y * gridDim.x);
}

template <int KIND, int ARITH, int RNG, int BC, int TH, int TJ>
static void launch_tile_t(const Geom &G, const Fields &F, int colour, const PhaseArgs &A, uint32_t *counts,
                          hipStream_t s)
{
    const uint32_t tiles_x = (G.half + TH - 1) / TH, tiles_y = (G.nj + TJ - 1) / TJ;
    hipLaunchKernelGGL((k_tile<KIND, ARITH, RNG, BC, TH, TJ>), dim3(tiles_x * tiles_y, G.nk), dim3(TH * TJ), 0, s, G,
                       F, colour, A, counts, tiles_x);
}

template <int KIND, int ARITH, int RNG, int BC, int TH, int TJ>
static void launch_march_t(const Geom &G, const Fields &F, int colour, const PhaseArgs &A, uint32_t *counts,
                           uint32_t nslab, hipStream_t s)
{
    const uint32_t tiles_x = (G.half + TH - 1) / TH, tiles_y = (G.nj + TJ - 1) / TJ;
    if (nslab < 1) nslab = 1;
    if (nslab > G.nk) nslab = G.nk;
    const uint32_t slab = (G.nk + nslab - 1) / nslab;
    const uint32_t ns = (G.nk + slab - 1) / slab;
    hipLaunchKernelGGL((k_march<KIND, ARITH, RNG, BC, TH, TJ>), dim3(tiles_x * tiles_y, ns), dim3(TH * TJ), 0, s, G, F,
                       colour, A, counts, tiles_x, slab);
}

template <int KIND, int ARITH, int RNG, int BC>
static int launch_march_tile(int tile, const Geom &G, const Fields &F, int colour, const PhaseArgs &A,
                             uint32_t *counts, uint32_t nslab, hipStream_t s)
{
    switch (tile) {
    case 0: launch_march_t<KIND, ARITH, RNG, BC, 64, 4>(G, F, colour, A, counts, nslab, s); return 0;
    case 1: launch_march_t<KIND, ARITH, RNG, BC, 32, 8>(G, F, colour, A, counts, nslab, s); return 0;
    case 2: launch_march_t<KIND, ARITH, RNG, BC, 16, 16>(G, F, colour, A, counts, nslab, s); return 0;
    case 3: launch_tile_t<KIND, ARITH, RNG, BC, 64, 4>(G, F, colour, A, counts, s); return 0;
    case 4: launch_tile_t<KIND, ARITH, RNG, BC, 32, 8>(G, F, colour, A, counts, s); return 0;
    case 5: launch_tile_t<KIND, ARITH, RNG, BC, 16, 16>(G, F, colour, A, counts, s); return 0;
    default: return -1;
    }
}

template <int KIND>
static int launch_march_kind(int arith, int rng, int bc, int tile, const Geom &G, const Fields &F, int colour,
                             const PhaseArgs &A, uint32_t *counts, uint32_t nslab, hipStream_t s)
{
    const bool ph = rng == RNG_PHILOX;
    if (arith == ARITH_COMPAT) {
        if (ph) return bc == BC_OPEN ? launch_march_tile<KIND, ARITH_COMPAT, RNG_PHILOX, BC_OPEN>(tile, G, F, colour, A, counts, nslab, s)
                                     : launch_march_tile<KIND, ARITH_COMPAT, RNG_PHILOX, BC_PERIODIC>(tile, G, F, colour, A, counts, nslab, s);
        return bc == BC_OPEN ? launch_march_tile<KIND, ARITH_COMPAT, RNG_STREAM, BC_OPEN>(tile, G, F, colour, A, counts, nslab, s)
                             : launch_march_tile<KIND, ARITH_COMPAT, RNG_STREAM, BC_PERIODIC>(tile, G, F, colour, A, counts, nslab, s);
    }
    if (ph) return bc == BC_OPEN ? launch_march_tile<KIND, ARITH_FP64, RNG_PHILOX, BC_OPEN>(tile, G, F, colour, A, counts, nslab, s)
                                 : launch_march_tile<KIND, ARITH_FP64, RNG_PHILOX, BC_PERIODIC>(tile, G, F, colour, A, counts, nslab, s);
    return bc == BC_OPEN ? launch_march_tile<KIND, ARITH_FP64, RNG_STREAM, BC_OPEN>(tile, G, F, colour, A, counts, nslab, s)
                         : launch_march_tile<KIND, ARITH_FP64, RNG_STREAM, BC_PERIODIC>(tile, G, F, colour, A, counts, nslab, s);
}

int launch_phase_march(int kind, int arith, int rng, int bc, int tile, const Geom &G, const Fields &F, int colour,
                       const PhaseArgs &A, uint32_t *counts, uint32_t nslab, hipStream_t s)
{
    return kind == 0 ? launch_march_kind<0>(arith, rng, bc, tile, G, F, colour, A, counts, nslab, s)
                     : launch_march_kind<1>(arith, rng, bc, tile, G, F, colour, A, counts, nslab, s);
}

uint32_t march_tiles(int tile, const Geom &G)
{
    static const uint32_t TH[6] = {64, 32, 16, 64, 32, 16}, TJ[6] = {4, 8, 16, 4, 8, 16};
    return ((G.half + TH[tile] - 1) / TH[tile]) * ((G.nj + TJ[tile] - 1) / TJ[tile]);
}

}  // namespace chll
