// chll_analysis.hip -- lattice analysis kernels (gfx950): shell-averaged
// nematic and chiral correlation functions and the chirality structure factor.
//
// Correlations (notebooks/paper_plots2.ipynb, cell 0: calc_n_corr and
// calc_s_corr).  For every site r and every offset d with |d|^2 < ns^2 whose
// partner r + d lies inside the lattice (open edges, no wrap), shell floor(|d|)
// collects the pair term and one count:
//   nematic  (n_r . n_{r+d})^2          g_n[sh] = sum / count
//   chiral   s_r s_{r+d} - <s>^2        g_s[sh] = sum / count
// Both pair terms are symmetric under d -> -d, so the kernel walks only the
// half space of offsets (dk > 0, or dk = 0 and dj > 0, or dk = dj = 0 and
// di > 0) and doubles; shell 0 holds d = 0 alone.  One thread per site, the
// offsets of one shell at a time (a wave-uniform table walk, so every partner
// load of a wave is one contiguous run), per-thread sums reduced per block in a
// fixed order and then per shell across blocks: results do not depend on the
// launch.  Chiral sums are exact integers; counts have the closed form
// sum_d (ni - |di|)(nj - |dj|)(nk - |dk|), evaluated on the host.
//
// Structure factor (notebooks/.ipynb_checkpoints/hero_stripe-checkpoint.py:
// 10-29, get_sc / structure_factor).  S(k) = |sum_r s_r exp(-i k . (r + 1))|^2
// / N^2 with k = kv * p for a plane vector p and a list of kv.  The phase
// depends on r only through the integer t = p . (r + 1), so one HBM pass over
// the spins builds the projection histogram h_p[t] = sum_{p.(r+1) = t} s_r for
// every plane at once (per-block LDS histograms, integer and exact), and a
// second kernel evaluates sum_t h_p[t] exp(-i kv t) for every (plane, kv):
// O(N) bytes plus O(planes x bins x T) phases instead of O(planes x bins x N).
//
// Layout: plain Fortran order (i fastest), directors as three float64 arrays,
// spins int8.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/chll_mod.h"
#include "chll_device.h"

namespace chll {

constexpr int kCorrBlock = 256;
constexpr int kSfBlock = 256;
constexpr int kSfHistBlock = 1024;
constexpr int kSfMaxBins = 12288;    // LDS histogram entries per pass (48 KiB)
constexpr int kSfMaxBlocks = 512;

struct CorrArgs {
    const double *nx, *ny, *nz;
    const int8_t *s;
    const int32_t *tab;      // packed half-space offsets, shell-major
    const int32_t *start;    // ns + 1 table offsets per shell
    double *pn;              // nblk x ns partial nematic sums
    long long *ps;           // nblk x (ns + 1) partial chiral sums, last column sum of s
    uint32_t ni, nj, nk, N, plane, nblk;
    int ns;
    FastDiv fd_ni, fd_nj;
};

__device__ inline double block_sum_d(double v, double *red)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kCorrBlock / 64; ++w) t += red[w];
    return t;
}

__device__ inline long long block_sum_l(long long v, long long *red)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    long long t = 0;
    for (int w = 0; w < kCorrBlock / 64; ++w) t += red[w];
    return t;
}

template <bool NEM, bool CHI>
__global__ __launch_bounds__(kCorrBlock) void k_shell_corr(CorrArgs A)
{
    __shared__ double redd[kCorrBlock / 64];
    __shared__ long long redl[kCorrBlock / 64];
    const uint32_t l = blockIdx.x * kCorrBlock + threadIdx.x;
    const bool live = l < A.N;
    int i = 0, j = 0, k = 0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    int sa = 0;
    if (live) {
        const uint32_t row = fdiv(l, A.fd_ni);
        i = (int)(l - row * A.ni);
        k = (int)fdiv(row, A.fd_nj);
        j = (int)row - k * (int)A.nj;
        if (NEM) { ax = A.nx[l]; ay = A.ny[l]; az = A.nz[l]; }
        if (CHI) sa = A.s[l];
    }
    for (int sh = 0; sh < A.ns; ++sh) {
        double accn = 0.0;
        long long accs = 0;
        if (sh == 0) {  // d = 0, counted once
            if (NEM) { const double c = (ax * ax + ay * ay) + az * az; accn = c * c; }
            if (CHI) accs = (long long)(sa * sa);
        } else {
            const int o1 = A.start[sh + 1];
            for (int o = A.start[sh]; o < o1; ++o) {
                const int32_t v = A.tab[o];
                const int di = (int)(int8_t)(v & 0xff), dj = (int)(int8_t)((v >> 8) & 0xff),
                          dk = (int)(int8_t)((v >> 16) & 0xff);
                const bool in = live && (uint32_t)(i + di) < A.ni && (uint32_t)(j + dj) < A.nj &&
                                (uint32_t)(k + dk) < A.nk;
                if (in) {
                    // modulo 2^32: exact for a partner inside the lattice
                    const uint32_t q = l + (uint32_t)di + (uint32_t)dj * A.ni + (uint32_t)dk * A.plane;
                    if (NEM) {
                        const double c = (ax * A.nx[q] + ay * A.ny[q]) + az * A.nz[q];
                        accn += c * c;
                    }
                    if (CHI) accs += (long long)(sa * (int)A.s[q]);
                }
            }
        }
        if (NEM) {
            const double t = block_sum_d(accn, redd);
            if (threadIdx.x == 0) A.pn[(size_t)blockIdx.x * A.ns + sh] = t;
        }
        if (CHI) {
            const long long t = block_sum_l(accs, redl);
            if (threadIdx.x == 0) A.ps[(size_t)blockIdx.x * (A.ns + 1) + sh] = t;
        }
    }
    if (CHI) {
        const long long t = block_sum_l((long long)sa, redl);
        if (threadIdx.x == 0) A.ps[(size_t)blockIdx.x * (A.ns + 1) + A.ns] = t;
    }
}

// One block per shell: fixed-order sums over the blocks, then the
// normalisation by the pair counts (shells above 0 doubled for the half space).
__global__ __launch_bounds__(kCorrBlock) void k_corr_final(CorrArgs A, const long long *counts, double *g_nem,
                                                          double *g_chi)
{
    __shared__ double redd[kCorrBlock / 64];
    __shared__ long long redl[kCorrBlock / 64];
    const int sh = blockIdx.x;
    const double cnt = (double)counts[sh];
    if (g_nem) {
        double v = 0.0;
        for (uint32_t b = threadIdx.x; b < A.nblk; b += kCorrBlock) v += A.pn[(size_t)b * A.ns + sh];
        const double t = block_sum_d(v, redd);
        if (threadIdx.x == 0) g_nem[sh] = (sh == 0 ? t : 2.0 * t) / cnt;
    }
    if (g_chi) {
        long long v = 0, m = 0;
        for (uint32_t b = threadIdx.x; b < A.nblk; b += kCorrBlock) {
            v += A.ps[(size_t)b * (A.ns + 1) + sh];
            m += A.ps[(size_t)b * (A.ns + 1) + A.ns];
        }
        const long long t = block_sum_l(v, redl);
        const long long tm = block_sum_l(m, redl);
        if (threadIdx.x == 0) {
            const double mean = (double)tm / (double)A.N;
            g_chi[sh] = (double)(sh == 0 ? t : 2 * t) / cnt - mean * mean;
        }
    }
}

struct SfPlane {
    int p0, p1, p2;
    int tmin;        // smallest p . (r + 1)
    int base;        // first histogram entry of this plane
    int T;           // histogram entries
};

constexpr int kSfMaxPlanes = 16;

struct SfArgs {
    const int8_t *s;
    int *part;       // nblk x ntot block histograms
    int *hist;       // ntot summed histograms
    uint32_t ni, nj, nk, N, rows, nblk;
    int nplanes, ntot;
    SfPlane pl[kSfMaxPlanes];
    FastDiv fd_nj;
};

// One wave per row (j, k) at a time, 16 waves of a block sharing one LDS
// histogram.  Each pass of the wave takes 64 U consecutive sites of the row:
// lane l holds sites ib + l + 64 u (U byte loads in flight), so every atomic
// round of a plane with |p0| = 1 hits 64 consecutive entries (no bank
// conflicts).  A plane with p0 != 0 sends each site to its own entry,
// p0 (i + 1) + c_row; a plane with p0 == 0 maps the whole row to one entry,
// which gets the wave's row sum.  Lane p holds plane p's constants (and, per
// row, its c_row); the plane loop broadcasts them with readlane, in
// wave-uniform control flow.
template <int U>
__global__ __launch_bounds__(kSfHistBlock) void k_sf_hist(SfArgs A)
{
    extern __shared__ int h[];
    for (int t = threadIdx.x; t < A.ntot; t += kSfHistBlock) h[t] = 0;
    __syncthreads();
    const int np = A.nplanes;
    const int lane = (int)(threadIdx.x & 63u);
    int q0 = 0, q1 = 0, q2 = 0, qb = 0;
    for (int p = 0; p < np; ++p) {
        if (lane == p) {
            q0 = A.pl[p].p0; q1 = A.pl[p].p1; q2 = A.pl[p].p2;
            qb = A.pl[p].base - A.pl[p].tmin;
        }
    }
    const uint32_t waves = A.nblk * (kSfHistBlock / 64);
    // the row is one per wave: keep it and its coordinates in scalar registers
    const uint32_t row0 =
        (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kSfHistBlock / 64) + (threadIdx.x >> 6)));
    for (uint32_t row = row0; row < A.rows; row += waves) {
        const uint32_t k = fdiv(row, A.fd_nj), j = row - k * A.nj;
        const int cst = q1 * (int)(j + 1) + q2 * (int)(k + 1) + qb;
        const int8_t *sr = A.s + (size_t)row * A.ni;
        int rsum = 0;
        for (uint32_t ib = 0; ib < A.ni; ib += 64u * U) {
            int v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = ib + 64u * u + (uint32_t)lane;
                v[u] = i < A.ni ? (int)sr[i] : 0;
                rsum += v[u];
            }
            for (int p = 0; p < np; ++p) {
                const int a = __builtin_amdgcn_readlane(q0, p);
                const int c = __builtin_amdgcn_readlane(cst, p);
                if (a != 0) {
                    const int t0 = a * (int)(ib + (uint32_t)lane + 1u) + c;
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (ib + 64u * u + (uint32_t)lane < A.ni) atomicAdd(&h[t0 + a * 64 * u], v[u]);
                }
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) rsum += __shfl_xor(rsum, off, 64);
        for (int p = 0; p < np; ++p) {
            const int a = __builtin_amdgcn_readlane(q0, p);
            const int c = __builtin_amdgcn_readlane(cst, p);
            if (a == 0 && lane == 0) atomicAdd(&h[c], rsum);
        }
    }
    __syncthreads();
    int *out = A.part + (size_t)blockIdx.x * A.ntot;
    for (int t = threadIdx.x; t < A.ntot; t += kSfHistBlock) out[t] = h[t];
}

// Block of 64 entries x 4 waves: wave w sums the block histograms w, w + 4,
// ... with four independent accumulators; the 4 partial sums are combined in
// LDS (integer sums: exact in any order).
__global__ __launch_bounds__(kSfBlock) void k_sf_reduce(SfArgs A)
{
    __shared__ int red[kSfBlock / 64][64];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const int t = blockIdx.x * 64 + (int)lane;
    int a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    if (t < A.ntot) {
        const size_t stride = (size_t)A.ntot;
        uint32_t b = w;
        for (; b + 12u < A.nblk; b += 16u) {
            a0 += A.part[(size_t)b * stride + t];
            a1 += A.part[(size_t)(b + 4u) * stride + t];
            a2 += A.part[(size_t)(b + 8u) * stride + t];
            a3 += A.part[(size_t)(b + 12u) * stride + t];
        }
        for (; b < A.nblk; b += 4u) a0 += A.part[(size_t)b * stride + t];
    }
    red[w][lane] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    if (w == 0 && t < A.ntot) A.hist[t] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// Block (bin, plane): sum_t h[t] exp(-i kv t) in a fixed order, then
// S = (Re / N)^2 + (Im / N)^2.
__global__ __launch_bounds__(kSfBlock) void k_sf_phase(SfArgs A, const double *kv, int bins, double *sf)
{
    __shared__ double red[2][kSfBlock / 64];
    const int b = blockIdx.x, p = blockIdx.y;
    const SfPlane P = A.pl[p];
    const double k = kv[b];
    double re = 0.0, im = 0.0;
    for (int u = threadIdx.x; u < P.T; u += kSfBlock) {
        const int hv = A.hist[P.base + u];
        if (hv != 0) {
            double sn, cs;
            sincos(k * (double)(P.tmin + u), &sn, &cs);
            re += (double)hv * cs;
            im -= (double)hv * sn;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
    }
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    if (lane == 0) { red[0][wid] = re; red[1][wid] = im; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0, ti = 0.0;
        for (int w = 0; w < kSfBlock / 64; ++w) { tr += red[0][w]; ti += red[1][w]; }
        const double n = (double)A.N;
        const double a = tr / n, c = ti / n;
        sf[(size_t)p * bins + b] = a * a + c * c;
    }
}

}  // namespace chll

using namespace chll;

namespace {
thread_local std::string g_aerr;
int afail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_aerr = buf;
    return -1;
}

// Host-to-device copies of small host tables without waiting for the stream:
// the tables go through one pinned staging buffer per host thread, which is
// reused only after the event recorded behind its previous copies has
// completed.  (The event belongs to the device current at the first call.)
struct HostStage {
    char *buf = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
};
thread_local HostStage g_stage;

// Copies bytes[i] bytes from host src[i] to device dst[i], i < k, in the
// order of stream st.
int stage_copy(int k, const void *const *src, void *const *dst, const size_t *bytes, hipStream_t st)
{
    HostStage &H = g_stage;
    size_t total = 0;
    for (int i = 0; i < k; ++i) total += (bytes[i] + 15) & ~(size_t)15;
    if (H.done && hipEventSynchronize(H.done) != hipSuccess) return afail("staging: event wait failed");
    if (!H.done && hipEventCreateWithFlags(&H.done, hipEventDisableTiming) != hipSuccess)
        return afail("staging: hipEventCreate failed");
    if (H.cap < total) {
        if (H.buf) (void)hipHostFree(H.buf);
        H.buf = nullptr;
        H.cap = 0;
        if (hipHostMalloc((void **)&H.buf, total) != hipSuccess) return afail("staging: hipHostMalloc(%zu) failed", total);
        H.cap = total;
    }
    size_t off = 0;
    for (int i = 0; i < k; ++i) {
        memcpy(H.buf + off, src[i], bytes[i]);
        if (hipMemcpyAsync(dst[i], H.buf + off, bytes[i], hipMemcpyHostToDevice, st) != hipSuccess)
            return afail("staging: copy failed");
        off += (bytes[i] + 15) & ~(size_t)15;
    }
    if (hipEventRecord(H.done, st) != hipSuccess) return afail("staging: hipEventRecord failed");
    return 0;
}

int isqrt_floor(int r2)
{
    int r = 0;
    while ((r + 1) * (r + 1) <= r2) ++r;
    return r;
}

bool half_space(int di, int dj, int dk) { return dk > 0 || (dk == 0 && (dj > 0 || (dj == 0 && di > 0))); }

// Half-space offsets with |d|^2 < ns^2, grouped by shell floor(|d|) (shell 0
// is empty: d = 0 is handled apart), in the order dk, dj, di within a shell.
void offset_table(int ns, std::vector<int32_t> &tab, std::vector<int32_t> &start)
{
    std::vector<std::vector<int32_t>> by(ns);
    const int m = ns - 1;
    for (int dk = -m; dk <= m; ++dk)
        for (int dj = -m; dj <= m; ++dj)
            for (int di = -m; di <= m; ++di) {
                const int r2 = di * di + dj * dj + dk * dk;
                if (r2 >= ns * ns || !half_space(di, dj, dk)) continue;
                const int32_t v = (int32_t)((uint32_t)(di & 0xff) | ((uint32_t)(dj & 0xff) << 8) |
                                            ((uint32_t)(dk & 0xff) << 16));
                by[isqrt_floor(r2)].push_back(v);
            }
    tab.clear();
    start.assign(ns + 1, 0);
    for (int sh = 0; sh < ns; ++sh) {
        start[sh] = (int32_t)tab.size();
        tab.insert(tab.end(), by[sh].begin(), by[sh].end());
    }
    start[ns] = (int32_t)tab.size();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct CorrLayout {
    size_t tab, start, counts, pn, ps, total;
    uint32_t nblk;
    size_t ntab;
};

CorrLayout corr_layout(int32_t ni, int32_t nj, int32_t nk, int32_t ns)
{
    CorrLayout L{};
    const uint64_t N = (uint64_t)ni * nj * nk;
    L.nblk = (uint32_t)((N + kCorrBlock - 1) / kCorrBlock);
    // half-space table size bound: (2 ns - 1)^3 / 2 + 1
    const uint64_t w = 2 * (uint64_t)ns - 1;
    L.ntab = (size_t)(w * w * w / 2 + 1);
    size_t off = 0;
    L.tab = off; off = align256(off + L.ntab * sizeof(int32_t));
    L.start = off; off = align256(off + (ns + 1) * sizeof(int32_t));
    L.counts = off; off = align256(off + ns * sizeof(long long));
    L.pn = off; off = align256(off + (size_t)L.nblk * ns * sizeof(double));
    L.ps = off; off = align256(off + (size_t)L.nblk * (ns + 1) * sizeof(long long));
    L.total = off;
    return L;
}

int corr_counts(int32_t ni, int32_t nj, int32_t nk, int32_t ns, int64_t *c)
{
    for (int sh = 0; sh < ns; ++sh) c[sh] = 0;
    const int m = ns - 1;
    for (int dk = -m; dk <= m; ++dk)
        for (int dj = -m; dj <= m; ++dj)
            for (int di = -m; di <= m; ++di) {
                const int r2 = di * di + dj * dj + dk * dk;
                if (r2 >= ns * ns) continue;
                const int64_t a = ni - (di < 0 ? -di : di), b = nj - (dj < 0 ? -dj : dj),
                              e = nk - (dk < 0 ? -dk : dk);
                if (a > 0 && b > 0 && e > 0) c[isqrt_floor(r2)] += a * b * e;
            }
    return 0;
}

struct SfLayout {
    size_t kv, part, hist, total;
    uint32_t nblk;
    int ntot;
};

// Plane table and histogram sizes; returns false (with the error set) on bad planes.
bool sf_planes(int32_t ni, int32_t nj, int32_t nk, const int32_t *planes, int32_t nplanes, SfPlane *pl, int &ntot)
{
    ntot = 0;
    const int64_t n[3] = {ni, nj, nk};
    for (int p = 0; p < nplanes; ++p) {
        const int32_t *v = planes + 3 * p;
        int64_t lo = 0, hi = 0;
        for (int c = 0; c < 3; ++c) {
            if (v[c] < -1024 || v[c] > 1024) { afail("chll_structure_factor: plane component %d out of range", v[c]); return false; }
            lo += v[c] > 0 ? (int64_t)v[c] : (int64_t)v[c] * n[c];
            hi += v[c] > 0 ? (int64_t)v[c] * n[c] : (int64_t)v[c];
        }
        const int64_t T = hi - lo + 1;
        if (ntot + T > kSfMaxBins) {
            afail("chll_structure_factor: planes need %lld histogram entries (limit %d per call)", (long long)(ntot + T),
                  kSfMaxBins);
            return false;
        }
        pl[p] = SfPlane{v[0], v[1], v[2], (int)lo, ntot, (int)T};
        ntot += (int)T;
    }
    return true;
}

SfLayout sf_layout(int32_t ni, int32_t nj, int32_t nk, int ntot, int32_t bins)
{
    SfLayout L{};
    const uint64_t rows = (uint64_t)nj * nk;
    const uint64_t want = (rows + kSfHistBlock / 64 - 1) / (kSfHistBlock / 64);
    L.nblk = (uint32_t)(want < kSfMaxBlocks ? want : kSfMaxBlocks);
    L.ntot = ntot;
    size_t off = 0;
    L.kv = off; off = align256(off + (size_t)bins * sizeof(double));
    L.part = off; off = align256(off + (size_t)L.nblk * ntot * sizeof(int));
    L.hist = off; off = align256(off + (size_t)ntot * sizeof(int));
    L.total = off;
    return L;
}
}  // namespace

extern "C" {

const char *chll_analysis_last_error(void) { return g_aerr.c_str(); }

int chll_corr_counts(int32_t ni, int32_t nj, int32_t nk, int32_t nshells, int64_t *counts)
{
    g_aerr.clear();
    if (!counts) return afail("chll_corr_counts: null argument");
    if (ni < 1 || nj < 1 || nk < 1) return afail("chll_corr_counts: bad shape");
    if (nshells < 1 || nshells > 127) return afail("chll_corr_counts: nshells must be in 1..127");
    return corr_counts(ni, nj, nk, nshells, counts);
}

size_t chll_corr_workspace_bytes(int32_t ni, int32_t nj, int32_t nk, int32_t nshells)
{
    if (ni < 1 || nj < 1 || nk < 1 || nshells < 1 || nshells > 127) return 0;
    return corr_layout(ni, nj, nk, nshells).total;
}

int chll_shell_corr(const double *nx, const double *ny, const double *nz, const int8_t *s, int32_t ni, int32_t nj,
                    int32_t nk, int32_t nshells, double *g_nem, double *g_chi, void *workspace, size_t workspace_bytes,
                    void *stream)
{
    g_aerr.clear();
    if (!g_nem && !g_chi) return afail("chll_shell_corr: no output requested");
    if (g_nem && (!nx || !ny || !nz)) return afail("chll_shell_corr: null director array");
    if (g_chi && !s) return afail("chll_shell_corr: null spin array");
    if (!workspace) return afail("chll_shell_corr: null workspace");
    if (ni < 1 || nj < 1 || nk < 1) return afail("chll_shell_corr: bad shape");
    if (nshells < 1 || nshells > 127) return afail("chll_shell_corr: nshells must be in 1..127");
    const uint64_t N = (uint64_t)ni * nj * nk;
    if (N >= (1ull << 31)) return afail("chll_shell_corr: lattice too large");
    const CorrLayout L = corr_layout(ni, nj, nk, nshells);
    if (workspace_bytes < L.total) return afail("chll_shell_corr: workspace too small");
    std::vector<int32_t> tab, start;
    offset_table(nshells, tab, start);
    std::vector<int64_t> cnt(nshells);
    corr_counts(ni, nj, nk, nshells, cnt.data());
    for (int sh = 0; sh < nshells; ++sh)
        if (cnt[sh] == 0) return afail("chll_shell_corr: shell %d has no pairs on a %dx%dx%d lattice", sh, ni, nj, nk);
    char *ws = (char *)workspace;
    hipStream_t st = (hipStream_t)stream;
    {
        const void *src[3] = {tab.data(), start.data(), cnt.data()};
        void *dst[3] = {ws + L.tab, ws + L.start, ws + L.counts};
        const size_t bytes[3] = {tab.size() * sizeof(int32_t), start.size() * sizeof(int32_t),
                                 cnt.size() * sizeof(int64_t)};
        if (stage_copy(3, src, dst, bytes, st)) return -1;
    }
    CorrArgs A;
    A.nx = nx; A.ny = ny; A.nz = nz; A.s = s;
    A.tab = (const int32_t *)(ws + L.tab);
    A.start = (const int32_t *)(ws + L.start);
    A.pn = (double *)(ws + L.pn);
    A.ps = (long long *)(ws + L.ps);
    A.ni = ni; A.nj = nj; A.nk = nk; A.N = (uint32_t)N; A.plane = (uint32_t)ni * nj; A.nblk = L.nblk;
    A.ns = nshells;
    A.fd_ni = make_fastdiv(ni); A.fd_nj = make_fastdiv(nj);
    const dim3 grid(L.nblk), block(kCorrBlock);
    if (g_nem && g_chi) hipLaunchKernelGGL((k_shell_corr<true, true>), grid, block, 0, st, A);
    else if (g_nem) hipLaunchKernelGGL((k_shell_corr<true, false>), grid, block, 0, st, A);
    else hipLaunchKernelGGL((k_shell_corr<false, true>), grid, block, 0, st, A);
    hipLaunchKernelGGL(k_corr_final, dim3(nshells), block, 0, st, A, (const long long *)(ws + L.counts), g_nem, g_chi);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return afail("chll_shell_corr launch: %s", hipGetErrorString(e));
    return 0;
}

size_t chll_sf_workspace_bytes(int32_t ni, int32_t nj, int32_t nk, const int32_t *planes, int32_t nplanes,
                               int32_t bins)
{
    if (ni < 1 || nj < 1 || nk < 1 || !planes || nplanes < 1 || nplanes > kSfMaxPlanes || bins < 1) return 0;
    SfPlane pl[kSfMaxPlanes];
    int ntot = 0;
    if (!sf_planes(ni, nj, nk, planes, nplanes, pl, ntot)) return 0;
    return sf_layout(ni, nj, nk, ntot, bins).total;
}

int chll_structure_factor(const int8_t *s, int32_t ni, int32_t nj, int32_t nk, const int32_t *planes,
                          int32_t nplanes, const double *kv, int32_t bins, double *sf, void *workspace,
                          size_t workspace_bytes, void *stream)
{
    g_aerr.clear();
    if (!s || !planes || !kv || !sf || !workspace) return afail("chll_structure_factor: null argument");
    if (ni < 1 || nj < 1 || nk < 1) return afail("chll_structure_factor: bad shape");
    if (nplanes < 1 || nplanes > kSfMaxPlanes) return afail("chll_structure_factor: nplanes must be in 1..%d", kSfMaxPlanes);
    if (bins < 1 || bins > 65535) return afail("chll_structure_factor: bins must be in 1..65535");
    const uint64_t N = (uint64_t)ni * nj * nk;
    if (N >= (1ull << 31)) return afail("chll_structure_factor: lattice too large");
    SfArgs A;
    int ntot = 0;
    if (!sf_planes(ni, nj, nk, planes, nplanes, A.pl, ntot)) return -1;
    const SfLayout L = sf_layout(ni, nj, nk, ntot, bins);
    if (workspace_bytes < L.total) return afail("chll_structure_factor: workspace too small");
    char *ws = (char *)workspace;
    hipStream_t st = (hipStream_t)stream;
    // kv is the caller's host array, staged so that the call does not wait for the stream
    {
        const void *src[1] = {kv};
        void *dst[1] = {ws + L.kv};
        const size_t bytes[1] = {(size_t)bins * sizeof(double)};
        if (stage_copy(1, src, dst, bytes, st)) return -1;
    }
    A.s = s;
    A.part = (int *)(ws + L.part);
    A.hist = (int *)(ws + L.hist);
    A.ni = ni; A.nj = nj; A.nk = nk; A.N = (uint32_t)N; A.rows = (uint32_t)((uint64_t)nj * nk); A.nblk = L.nblk;
    A.nplanes = nplanes; A.ntot = ntot;
    A.fd_nj = make_fastdiv(nj);
    hipLaunchKernelGGL(k_sf_hist<4>, dim3(L.nblk), dim3(kSfHistBlock), (size_t)ntot * sizeof(int), st, A);
    hipLaunchKernelGGL(k_sf_reduce, dim3((ntot + 63) / 64), dim3(kSfBlock), 0, st, A);
    hipLaunchKernelGGL(k_sf_phase, dim3(bins, nplanes), dim3(kSfBlock), 0, st, A, (const double *)(ws + L.kv), bins, sf);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return afail("chll_structure_factor launch: %s", hipGetErrorString(e));
    return 0;
}

}  // extern "C"
