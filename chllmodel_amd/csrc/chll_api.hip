// chll_api.hip -- C ABI of libchll_mod.so: the drop-in init/run entry points
// of chll/chll_mod.f90 and the device-resident lattice extensions.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/chll_mod.h"
#include "chll_kernels.h"

using namespace chll;

namespace {

thread_local std::string g_err;

int fail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return -1;
}

#define HIPCHECK(x)                                                                   \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) return fail("%s: %s", #x, hipGetErrorString(e_));       \
    } while (0)

template <class T>
int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) return fail("hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    return 0;
}

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// Options of the drop-in run() (set by chll_configure / chll_set_*).
struct RunConfig {
    int rng = RNG_PHILOX, arith = ARITH_COMPAT, bc = BC_OPEN, verbose = 1;
    uint64_t seed = 0x43484C4C2D4D4335ull;  // "CHLL-MC5"
    uint32_t run_index = 0;
    uint64_t stream_seed = 0, stream_pos = 0;
    const double *buffer = nullptr;
    int64_t buffer_len = 0;
    std::vector<double> last_rows;
};

std::mutex g_mu;
RunConfig g_cfg;
bool g_cfg_env_read = false;

}  // namespace

enum {
    KIND_DIR = 0, KIND_SPIN = 1, KIND_MEASURE = 2, KIND_FINAL = 3, KIND_FILL = 4, KIND_FUSED = 5, KIND_FUSED_NEXT = 6,
    KIND_HELD = 7, KIND_COUNT = 8
};

struct chll_lattice {
    Geom G{};
    int bc = BC_OPEN;
    int device = 0;  // HIP device current at creation
    hipStream_t own = nullptr, stream = nullptr;
    Fields F{};
    double *stage = nullptr;
    int32_t *stage_s = nullptr;
    int *bad = nullptr;
    uint32_t *counts = nullptr;
    uint32_t nblk = 0, nblk_meas = 0;
    double *part = nullptr;
    long long *spart = nullptr;
    double *rows_dev = nullptr;
    int32_t rows_cap = 0;
    double *d_dev = nullptr;
    long long *pend = nullptr;  // device: 2 director accept totals awaiting their flips
    double *ustream = nullptr;
    long long *counts_dev = nullptr;
    size_t counts_cap = 0;
    double *pinned = nullptr;  // host: 8 doubles
    // half-sweep kernel per [kind: director, chirality][arith: compat, fp64]:
    // variant 0 = one site per thread (tile bit 0: XCD row bands, bit 1: two
    // sites per thread with pair loads), 1 = k-march through LDS (tile shape,
    // slabs)
    struct KCfg { int variant, tile; uint32_t nslab; } kcfg[2][2];
    // FP64 sweeps: run each colour-1 site's chirality trial inside the colour-1
    // director launch (needs the one-site director kernel)
    bool fuse = true;
    double screen = kScreenScale;  // compat_screen's scale for reference-arithmetic sweeps (0: off)
    // reference-arithmetic fused sweeps (held chirality trials): flip bitmap,
    // queue of held trials, and the count of trials held since creation
    unsigned long long *flipbits = nullptr;
    uint32_t *held = nullptr;
    unsigned long long *held_total = nullptr;
    bool timing = false;
    struct Ev { int kind; hipEvent_t a, b; };
    std::vector<Ev> ev;
    size_t ev_used = 0;
};

namespace {

int check_geometry(int32_t ni, int32_t nj, int32_t nk, int32_t bc)
{
    if (ni < 1 || nj < 1 || nk < 1) return fail("lattice dims must be positive (got %d,%d,%d)", ni, nj, nk);
    const uint64_t n3 = (uint64_t)ni * nj * nk;
    // the reference's init writes ceil(N/2) entries into lists of N/2 (chll/chll_mod.f90:23-36, chll/chll.py:54)
    if (n3 % 2) return fail("N = ni*nj*nk must be even (got %d,%d,%d)", ni, nj, nk);
    if (n3 >= (1ull << 32)) return fail("lattice too large: N = %llu sites (limit 2^32)", (unsigned long long)n3);
    if (bc != BC_OPEN && bc != BC_PERIODIC) return fail("unknown boundary condition %d", bc);
    if (bc == BC_PERIODIC && ((ni % 2) || (nj % 2) || (nk % 2)))
        return fail("periodic boundaries need even ni, nj and nk for a valid checkerboard (got %d,%d,%d)", ni, nj, nk);
    return 0;
}

int tbegin(chll_lattice *L, int kind, size_t *idx)
{
    if (!L->timing) return 0;
    if (L->ev_used == L->ev.size()) {
        chll_lattice::Ev e{kind, nullptr, nullptr};
        HIPCHECK(hipEventCreate(&e.a));
        HIPCHECK(hipEventCreate(&e.b));
        L->ev.push_back(e);
    }
    *idx = L->ev_used++;
    L->ev[*idx].kind = kind;
    HIPCHECK(hipEventRecord(L->ev[*idx].a, L->stream));
    return 0;
}

int tend(chll_lattice *L, size_t idx)
{
    if (!L->timing) return 0;
    HIPCHECK(hipEventRecord(L->ev[idx].b, L->stream));
    return 0;
}

int launch_check()
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail("kernel launch: %s", hipGetErrorString(e));
    return 0;
}

}  // namespace

extern "C" {

const char *chll_last_error(void) { return g_err.c_str(); }

int chll_lattice_create(int32_t ni, int32_t nj, int32_t nk, int32_t bc, chll_lattice **out)
{
    if (!out) return fail("chll_lattice_create: null output pointer");
    *out = nullptr;
    if (check_geometry(ni, nj, nk, bc)) return -1;
    if ((uint64_t)ni * nj * nk + kFieldPad >= (1ull << 32)) return fail("lattice too large for 32-bit site indices");
    chll_lattice *L = new chll_lattice();
    (void)hipGetDevice(&L->device);
    Geom &G = L->G;
    G.ni = (uint32_t)ni; G.nj = (uint32_t)nj; G.nk = (uint32_t)nk;
    G.half = G.ni / 2;
    G.plane = G.half * G.nj;
    G.nsub = (uint32_t)(((uint64_t)ni * nj * nk) / 2);
    G.zero = 2 * G.nsub;
    G.nplane = G.ni * G.nj;
    G.fd_half = make_fastdiv(G.half ? G.half : 1u);
    G.fd_nj = make_fastdiv(G.nj);
    G.fd_ni = make_fastdiv(G.ni);
    L->bc = bc;
    const size_t ns = G.nsub, n3 = 2 * ns;
    L->nblk = phase_blocks(G);
    int rc = 0;
    // fields carry kFieldPad zeroed elements after the 2 nsub sites (the zero site)
    rc |= dalloc(&L->F.nx, n3 + kFieldPad);
    rc |= dalloc(&L->F.ny, n3 + kFieldPad);
    rc |= dalloc(&L->F.nz, n3 + kFieldPad);
    rc |= dalloc(&L->F.s, n3 + kFieldPad);
    if (!rc && (hipMemset(L->F.nx + n3, 0, kFieldPad * sizeof(double)) != hipSuccess ||
                hipMemset(L->F.ny + n3, 0, kFieldPad * sizeof(double)) != hipSuccess ||
                hipMemset(L->F.nz + n3, 0, kFieldPad * sizeof(double)) != hipSuccess ||
                hipMemset(L->F.s + n3, 0, kFieldPad * sizeof(int8_t)) != hipSuccess))
        rc = fail("hipMemset of the zero-site padding failed");
    rc |= dalloc(&L->bad, 1);
    L->nblk_meas = measure_blocks(G);
    rc |= dalloc(&L->counts, (size_t)kCountWords);
    rc |= dalloc(&L->part, 8 * (size_t)L->nblk_meas);
    rc |= dalloc(&L->spart, (size_t)L->nblk_meas);
    if (!rc && hipMemset(L->counts, 0, kCountWords * sizeof(uint32_t)) != hipSuccess) rc = fail("hipMemset failed");
    rc |= dalloc(&L->d_dev, 1);
    rc |= dalloc(&L->pend, 2);
    if (!rc && hipHostMalloc((void **)&L->pinned, 8 * sizeof(double)) != hipSuccess) rc = fail("hipHostMalloc failed");
    if (!rc && hipStreamCreate(&L->own) != hipSuccess) rc = fail("hipStreamCreate failed");
    if (rc) {
        std::string keep = g_err;
        chll_lattice_destroy(L);
        g_err = keep;
        return -1;
    }
    L->stream = L->own;
    // Default kernels (MI355X measurements at L = 256 and 128): site pairs in
    // XCD row bands where the rows allow them, else one site per thread for
    // the director trials (a carried plane ring would halve their register
    // budget) and the k-march for the chirality trials.  March shape: the
    // widest tile the rows fill, up to ~2048 blocks with slabs of at least 8
    // planes; lattices that cannot fill 512 such blocks keep the one-site
    // kernels.
    {
        const int tile = G.half >= 64 ? 0 : (G.half >= 32 ? 1 : 2);
        const uint32_t tiles = march_tiles(tile, G);
        uint32_t ns = (2048u + tiles - 1) / tiles;
        const uint32_t ns8 = G.nk / 8 ? G.nk / 8 : 1;
        ns = ns < ns8 ? ns : ns8;
        const bool thick = (G.nk + ns - 1) / ns >= 8 && tiles * ns >= 512;
        for (int kind = 0; kind < 2; ++kind)
            for (int ar = 0; ar < 2; ++ar) {
                // FP64: two sites per thread in XCD row bands where the rows
                // allow it (2-5 % faster fused sweeps), the fused colour-0
                // launch in reverse block order (it starts on the planes the
                // colour-1 launch finished with: 5 % faster sweeps at L = 256);
                // the chirality launches outside the fused chain (a call's
                // last sweep, measurement sweeps) on site pairs too, which
                // read the state once where the k-march re-reads its slab
                // halos (1.6x).  Compat: site pairs in bands for both kinds
                // (5 % faster sweeps at L = 256 than the k-march chirality),
                // which its fused sweeps need (compat_fused).
                // Without pairs: one-site kernels in bands and the k-march.
                const bool pairs = pairs_ok(G);
                const bool march = !pairs && thick && kind == 1;
                const int order = pairs ? 7 : 1;
                L->kcfg[kind][ar] = {march ? 1 : 0, march ? tile : order, ns};
                // odd ni: rows hold (ni +- 1) / 2 sites of a colour, so only the
                // one-site kernels in lattice order apply
                if (G.ni & 1u) L->kcfg[kind][ar] = {0, 0, 1};
            }
    }
    *out = L;
    return 0;
}

int chll_lattice_set_kernel(chll_lattice *L, int32_t kind, int32_t arith, int32_t variant, int32_t tile,
                            int32_t nslab)
{
    if (!L) return fail("null lattice");
    if (kind != 0 && kind != 1) return fail("kind must be 0 (director) or 1 (chirality)");
    if (arith != ARITH_COMPAT && arith != ARITH_FP64) return fail("unknown arithmetic mode %d", arith);
    if (variant != 0 && variant != 1) return fail("kernel variant must be 0 (site) or 1 (march)");
    if (variant == 0 && (tile < 0 || tile > 7))
        return fail("site kernel order must be 0..7 (bit 0: XCD row bands, bit 1: site pairs, "
                    "bit 2: fused colour-0 pair launch in reverse block order)");
    if (variant == 1 && (tile < 0 || tile > 5))
        return fail("tile must be 0..2 (k-march 64x4, 32x8, 16x16) or 3..5 (plane tiles)");
    if (nslab < 1) return fail("nslab must be >= 1");
    if ((L->G.ni & 1u) && (variant != 0 || tile != 0))
        return fail("odd ni (%u) takes only the one-site kernel in lattice order (variant 0, tile 0)", L->G.ni);
    L->kcfg[kind][arith] = {variant, tile, (uint32_t)nslab > L->G.nk ? L->G.nk : (uint32_t)nslab};
    return 0;
}

int chll_lattice_set_fusion(chll_lattice *L, int32_t on)
{
    if (!L) return fail("null lattice");
    L->fuse = on != 0;
    return 0;
}

int chll_lattice_get_fusion(chll_lattice *L, int32_t *on)
{
    if (!L || !on) return fail("null argument");
    *on = L->fuse ? 1 : 0;
    return 0;
}

int chll_lattice_set_screen(chll_lattice *L, double scale)
{
    if (!L) return fail("null lattice");
    if (!(scale == 0.0 || (scale >= kScreenScale && scale <= 1.0)))
        return fail("screen scale must be 0 (off) or in [2^-21, 1] (got %g)", scale);
    L->screen = scale;
    return 0;
}

int chll_lattice_get_kernel(chll_lattice *L, int32_t kind, int32_t arith, int32_t *variant, int32_t *tile,
                            int32_t *nslab)
{
    if (!L) return fail("null lattice");
    if ((kind != 0 && kind != 1) || (arith != ARITH_COMPAT && arith != ARITH_FP64)) return fail("bad kind/arith");
    const auto &c = L->kcfg[kind][arith];
    if (variant) *variant = c.variant;
    if (tile) *tile = c.tile;
    if (nslab) *nslab = (int32_t)c.nslab;
    return 0;
}

int chll_lattice_destroy(chll_lattice *L)
{
    if (!L) return 0;
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    dfree(L->F.nx); dfree(L->F.ny); dfree(L->F.nz); dfree(L->F.s);
    dfree(L->stage); dfree(L->stage_s); dfree(L->bad); dfree(L->counts); dfree(L->pend);
    dfree(L->part); dfree(L->spart); dfree(L->rows_dev); dfree(L->d_dev);
    dfree(L->ustream); dfree(L->counts_dev);
    dfree(L->flipbits); dfree(L->held); dfree(L->held_total);
    if (L->pinned) (void)hipHostFree(L->pinned);
    for (auto &e : L->ev) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    if (L->own) (void)hipStreamDestroy(L->own);
    delete L;
    return 0;
}

int chll_lattice_set_stream(chll_lattice *L, void *s)
{
    if (!L) return fail("null lattice");
    L->stream = s ? (hipStream_t)s : L->own;
    return 0;
}

static int ensure_stage(chll_lattice *L)
{
    const size_t n3 = 2 * (size_t)L->G.nsub;
    if (!L->stage && dalloc(&L->stage, n3)) return -1;
    if (!L->stage_s && dalloc(&L->stage_s, n3)) return -1;
    return 0;
}

// The flip bitmap and the queue of held trials (one slot per colour-1 site),
// allocated on a lattice's first fused reference-arithmetic sweep.
static int ensure_held(chll_lattice *L)
{
    if (L->held) return 0;
    if (!L->flipbits && dalloc(&L->flipbits, flipbit_words(L->G))) return -1;
    if (!L->held_total && dalloc(&L->held_total, 1)) return -1;
    if (dalloc(&L->held, kHeldHead + (size_t)L->G.nsub)) return -1;
    if (hipMemset(L->held, 0, kHeldHead * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(L->held_total, 0, sizeof(unsigned long long)) != hipSuccess) {
        dfree(L->held);
        return fail("hipMemset of the held-trial queue failed");
    }
    return 0;
}

int chll_lattice_upload(chll_lattice *L, const double *nx, const double *ny, const double *nz, const int32_t *s)
{
    if (!L || !nx || !ny || !nz || !s) return fail("chll_lattice_upload: null argument");
    if (ensure_stage(L)) return -1;
    const Geom &G = L->G;
    const size_t n3 = 2 * (size_t)G.nsub;
    const double *src[3] = {nx, ny, nz};
    double *dst[3] = {L->F.nx, L->F.ny, L->F.nz};
    for (int c = 0; c < 3; ++c) {
        HIPCHECK(hipMemcpyAsync(L->stage, src[c], n3 * sizeof(double), hipMemcpyHostToDevice, L->stream));
        launch_scatter_f64(G, L->stage, dst[c], L->stream);
        if (launch_check()) return -1;
    }
    HIPCHECK(hipMemsetAsync(L->bad, 0, sizeof(int), L->stream));
    HIPCHECK(hipMemcpyAsync(L->stage_s, s, n3 * sizeof(int32_t), hipMemcpyHostToDevice, L->stream));
    launch_scatter_spin(G, L->stage_s, L->F.s, L->bad, L->stream);
    if (launch_check()) return -1;
    int bad = 0;
    HIPCHECK(hipMemcpyAsync(&L->pinned[0], L->bad, sizeof(int), hipMemcpyDeviceToHost, L->stream));
    HIPCHECK(hipStreamSynchronize(L->stream));
    memcpy(&bad, &L->pinned[0], sizeof(int));
    if (bad) return fail("spin values must lie in [-127, 127]");
    return 0;
}

int chll_lattice_download(chll_lattice *L, double *nx, double *ny, double *nz, int32_t *s)
{
    if (!L || !nx || !ny || !nz || !s) return fail("chll_lattice_download: null argument");
    if (ensure_stage(L)) return -1;
    const Geom &G = L->G;
    const size_t n3 = 2 * (size_t)G.nsub;
    double *dst[3] = {nx, ny, nz};
    const double *src[3] = {L->F.nx, L->F.ny, L->F.nz};
    for (int c = 0; c < 3; ++c) {
        launch_gather_f64(G, src[c], L->stage, L->stream);
        if (launch_check()) return -1;
        HIPCHECK(hipMemcpyAsync(dst[c], L->stage, n3 * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    }
    launch_gather_spin(G, L->F.s, L->stage_s, L->stream);
    if (launch_check()) return -1;
    HIPCHECK(hipMemcpyAsync(s, L->stage_s, n3 * sizeof(int32_t), hipMemcpyDeviceToHost, L->stream));
    HIPCHECK(hipStreamSynchronize(L->stream));
    return 0;
}

int chll_lattice_init_state(chll_lattice *L, int32_t kind, uint64_t seed, uint32_t replica, double p_flip)
{
    if (!L) return fail("null lattice");
    if (kind < CHLL_INIT_ISOTROPIC || kind > CHLL_INIT_ALIGNED_XZ) return fail("unknown initial state kind %d", kind);
    if (!(p_flip >= 0.0 && p_flip <= 1.0)) return fail("p_flip must lie in [0, 1]");
    launch_init_state(L->G, L->F, kind, seed, replica, p_flip, L->stream);
    if (launch_check()) return -1;
    HIPCHECK(hipStreamSynchronize(L->stream));
    return 0;
}

int chll_lattice_set_step(chll_lattice *L, double d)
{
    if (!L) return fail("null lattice");
    HIPCHECK(hipStreamSynchronize(L->stream));
    L->pinned[1] = d;
    HIPCHECK(hipMemcpyAsync(L->d_dev, &L->pinned[1], sizeof(double), hipMemcpyHostToDevice, L->stream));
    HIPCHECK(hipStreamSynchronize(L->stream));
    return 0;
}

int chll_lattice_get_step(chll_lattice *L, double *d)
{
    if (!L || !d) return fail("null argument");
    HIPCHECK(hipMemcpyAsync(&L->pinned[2], L->d_dev, sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIPCHECK(hipStreamSynchronize(L->stream));
    *d = L->pinned[2];
    return 0;
}

int chll_lattice_sync(chll_lattice *L)
{
    if (!L) return fail("null lattice");
    HIPCHECK(hipStreamSynchronize(L->stream));
    return 0;
}

int chll_lattice_set_timing(chll_lattice *L, int32_t on)
{
    if (!L) return fail("null lattice");
    HIPCHECK(hipStreamSynchronize(L->stream));
    L->timing = on != 0;
    L->ev_used = 0;
    return 0;
}

int chll_lattice_held_trials(chll_lattice *L, int64_t *n)
{
    if (!L || !n) return fail("null argument");
    *n = 0;
    if (!L->held_total) return 0;
    unsigned long long v = 0;
    HIPCHECK(hipMemcpyAsync(&v, L->held_total, sizeof(v), hipMemcpyDeviceToHost, L->stream));
    HIPCHECK(hipStreamSynchronize(L->stream));
    *n = (int64_t)v;
    return 0;
}

int chll_lattice_timing(chll_lattice *L, double *ms, int64_t *launches, int32_t nkinds)
{
    if (!L) return fail("null lattice");
    HIPCHECK(hipStreamSynchronize(L->stream));
    for (int k = 0; k < nkinds; ++k) { if (ms) ms[k] = 0.0; if (launches) launches[k] = 0; }
    for (size_t i = 0; i < L->ev_used; ++i) {
        float t = 0.f;
        HIPCHECK(hipEventElapsedTime(&t, L->ev[i].a, L->ev[i].b));
        const int k = L->ev[i].kind;
        if (k < nkinds) {
            if (ms) ms[k] += t;
            if (launches) launches[k] += 1;
        }
    }
    L->ev_used = 0;
    return 0;
}

int chll_bandwidth_probe(int32_t pattern, int64_t bytes, int32_t reps, double *best_gbps, double *median_gbps)
{
    if (pattern < CHLL_BW_COPY || pattern > CHLL_BW_WRITE) return fail("unknown bandwidth pattern %d", pattern);
    if (bytes < (256ll << 20) || bytes > (16ll << 30)) return fail("probe arrays must hold 256 MiB to 16 GiB each");
    if (reps < 1 || reps > 1000) return fail("reps must lie in 1..1000");
    const uint64_t nb = (uint64_t)bytes & ~(uint64_t)15;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<float> ms;
    int rc = 0;
    const auto step = [&](hipError_t e, const char *what) {
        if (!rc && e != hipSuccess) rc = fail("chll_bandwidth_probe: %s: %s", what, hipGetErrorString(e));
    };
    step(hipMalloc(&a, nb), "hipMalloc");
    step(hipMalloc(&b, nb), "hipMalloc");
    step(hipStreamCreate(&st), "hipStreamCreate");
    step(hipEventCreate(&e0), "hipEventCreate");
    step(hipEventCreate(&e1), "hipEventCreate");
    if (!rc) step(hipMemsetAsync(a, 0, nb, st), "hipMemsetAsync");
    if (!rc) step(hipMemsetAsync(b, 0, nb, st), "hipMemsetAsync");
    for (int r = 0; !rc && r < 3 + reps; ++r) {
        if (r >= 3) step(hipEventRecord(e0, st), "hipEventRecord");
        if (!rc && launch_bandwidth(pattern, a, b, nb, st)) rc = fail("chll_bandwidth_probe: bad launch size");
        if (!rc) step(hipGetLastError(), "kernel launch");
        if (r < 3) continue;
        step(hipEventRecord(e1, st), "hipEventRecord");
        step(hipEventSynchronize(e1), "hipEventSynchronize");
        float t = 0.f;
        step(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
        ms.push_back(t);
    }
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (rc) return -1;
    std::sort(ms.begin(), ms.end());
    const double moved = (double)nb * (pattern == CHLL_BW_COPY ? 2.0 : pattern == CHLL_BW_READ2_WRITE1 ? 3.0 : 1.0);
    if (best_gbps) *best_gbps = moved / (ms.front() * 1e-3) / 1e9;
    if (median_gbps) *median_gbps = moved / (ms[ms.size() / 2] * 1e-3) / 1e9;
    return 0;
}

static int ensure_rows(chll_lattice *L, int32_t rows)
{
    if (rows <= L->rows_cap) return 0;
    dfree(L->rows_dev);
    L->rows_cap = 0;
    if (dalloc(&L->rows_dev, (size_t)rows * kRowCols)) return -1;
    L->rows_cap = rows;
    return 0;
}

static bool fused_sweep(const chll_lattice *L, int arith)
{
    return arith == ARITH_FP64 && L->fuse && L->kcfg[0][ARITH_FP64].variant == 0;
}

// Reference arithmetic fuses colour 1's chirality trials into its director
// launch through held trials (chll_site.h): it needs the screen, a positive
// kT at every sweep (the screen decides nothing at kT <= 0), and site pairs
// for the director and chirality kernels.
static bool compat_fused(const chll_lattice *L, const chll_params *p)
{
    const auto &kd = L->kcfg[0][ARITH_COMPAT], &ks = L->kcfg[1][ARITH_COMPAT];
    return p->arith == ARITH_COMPAT && L->fuse && L->screen > 0.0 && p->kbt > 0.0 && kd.variant == 0 &&
           ks.variant == 0 && compat_fused_ok(L->G, kd.tile, ks.tile);
}

int chll_lattice_sweeps(chll_lattice *L, chll_params *p, int32_t nstart, int32_t nstop, int32_t nout,
                        double *rows, int32_t max_rows, int64_t *counts)
{
    if (!L || !p) return fail("chll_lattice_sweeps: null argument");
    if (p->rng != RNG_PHILOX && p->rng != RNG_STREAM && p->rng != RNG_BUFFER) return fail("unknown rng mode %d", p->rng);
    if (p->arith != ARITH_COMPAT && p->arith != ARITH_FP64) return fail("unknown arithmetic mode %d", p->arith);
    if (!(p->kbt_end >= 0.0 && p->kbt_end < HUGE_VAL))
        return fail("kbt_end must be 0 (no ramp) or a finite positive end temperature (got %g)", p->kbt_end);
    if (max_rows < 0) max_rows = 0;
    const int64_t nsw = (int64_t)nstop - nstart + 1;
    if (nsw <= 0) return 0;
    const Geom &G = L->G;
    const uint64_t ns = G.nsub;
    if (p->rng == RNG_BUFFER) {
        if (!p->buffer) return fail("rng=buffer without a uniform buffer");
        if ((int64_t)(p->stream_pos + (uint64_t)nsw * 10 * ns) > p->buffer_len)
            return fail("uniform buffer too short: %lld sweeps need %llu draws from position %llu, buffer holds %lld",
                        (long long)nsw, (unsigned long long)(nsw * 10 * ns), (unsigned long long)p->stream_pos,
                        (long long)p->buffer_len);
    }
    if (p->rng != RNG_PHILOX && !L->ustream && dalloc(&L->ustream, 6 * ns)) return -1;
    if (ensure_rows(L, max_rows)) return -1;
    if (counts && (size_t)nsw * 2 > L->counts_cap) {
        dfree(L->counts_dev);
        L->counts_cap = 0;
        if (dalloc(&L->counts_dev, (size_t)nsw * 2)) return -1;
        L->counts_cap = (size_t)nsw * 2;
    }
    hipStream_t st = L->stream;
    // the temperature of sweep istep: kbt, or the linear ramp from kbt at
    // nstart to kbt_end at nstop (the bracketing and operation order are part
    // of the interface: callers reproduce the values in binary64)
    const bool ramp = p->kbt_end > 0.0 && nstop > nstart;
    const auto kt_of = [&](int64_t istep) {
        if (!ramp) return p->kbt;
        if (istep == nstop) return p->kbt_end;
        return p->kbt + (p->kbt_end - p->kbt) * (double)(istep - nstart) / (double)(nstop - nstart);
    };
    PhaseArgs A{};
    A.g = p->g; A.KK = p->KK; A.kbt = p->kbt; A.inv_kbt = 1.0 / p->kbt; A.inv_kbt_spin = A.inv_kbt;
    A.kbt_spin = A.kbt;
    A.d = L->d_dev; A.seed = p->seed; A.run_index = p->run_index; A.replica = p->replica;
    A.screen = L->screen;
    FinArgs FA{};
    FA.counts = L->counts; FA.pend = L->pend; FA.part = L->part; FA.spart = L->spart;
    FA.nblk_meas = L->nblk_meas; FA.n3 = 2 * ns; FA.d = L->d_dev; FA.rows = L->rows_dev;
    FA.counts_out = counts ? L->counts_dev : nullptr; FA.adapt = p->adapt;
    int32_t row = 0;
    size_t ti = 0;
    const int rngk = p->rng == RNG_PHILOX ? RNG_PHILOX : RNG_STREAM;
    // Sweep schedule.  Reference order per sweep: director colour 0 (D0),
    // director colour 1 (D1), chirality colour 0 (C0), chirality colour 1 (C1),
    // then the finalize kernel.  Fused: C1 runs inside the D1 launch (in
    // reference arithmetic with its open trials held and decided before C0
    // runs); with Philox draws also C0 of sweep t and D0 of sweep t + 1 share
    // a launch, after a finalize that closes sweep t's director accepts (and
    // so its cone-step update).  Sweep t's flips are then closed by the next
    // finalize.  Measurement sweeps and the last sweep of the call run their
    // C0 alone, so the state they leave is complete.
    const bool cfuse = compat_fused(L, p);
    const bool fuse1 = fused_sweep(L, p->arith) || cfuse;
    const bool fuse2 = fuse1 && p->rng == RNG_PHILOX;
    if (cfuse) {
        if (ensure_held(L)) return -1;
        launch_held_clear(L->held, st);
        if (launch_check()) return -1;
    }
    const auto grp = [&](int g) { return L->counts + (size_t)g * kSlots * kSlotStride; };
    bool d0_done = false, prev_open = false;
    for (int64_t t = 0; t < nsw; ++t) {
        const int32_t istep = nstart + (int32_t)t;
        if (p->rng == RNG_STREAM) {
            if (tbegin(L, KIND_FILL, &ti)) return -1;
            launch_stream_fill(L->ustream, p->seed, p->stream_pos, G.nsub, st);
            if (launch_check() || tend(L, ti)) return -1;
            p->stream_pos += 10 * ns;
        } else if (p->rng == RNG_BUFFER) {
            static const uint64_t off[6] = {0, 1, 3, 4, 7, 9};
            for (int q = 0; q < 6; ++q)
                HIPCHECK(hipMemcpyAsync(L->ustream + q * ns, p->buffer + p->stream_pos + off[q] * ns,
                                        ns * sizeof(double), hipMemcpyHostToDevice, st));
            p->stream_pos += 10 * ns;
        }
        A.istep = (uint32_t)istep;
        if (ramp) {
            A.kbt = kt_of(istep);
            A.inv_kbt = 1.0 / A.kbt;
            A.inv_kbt_spin = A.inv_kbt;
            A.kbt_spin = A.kbt;
        }
        const auto set_u = [&](int kind, int colour) {
            if (p->rng == RNG_PHILOX) return;
            if (kind == 0) { A.u_phi = L->ustream + (2 * colour) * ns; A.u_acc = L->ustream + (2 * colour + 1) * ns; }
            else { A.u_phi = nullptr; A.u_acc = L->ustream + (4 + colour) * ns; }
            A.u_spin = L->ustream + (4 + colour) * ns;
        };
        const auto half = [&](int kind, int colour, uint32_t *slots) -> int {
            set_u(kind, colour);
            const auto &kc = L->kcfg[kind][p->arith];
            if (tbegin(L, kind == 0 ? KIND_DIR : KIND_SPIN, &ti)) return -1;
            if (kc.variant == 1) {
                if (launch_phase_march(kind, p->arith, rngk, L->bc, kc.tile, G, L->F, colour, A, slots, kc.nslab, st))
                    return fail("unknown march tile %d", kc.tile);
            } else {
                // colour-1 half-sweeps walk their blocks in reverse, so every
                // launch starts where the one before it finished
                launch_phase(kind, p->arith, rngk, L->bc, G, L->F, colour, A, slots, (kc.tile & 3) | (colour << 2),
                             st);
            }
            return (launch_check() || tend(L, ti)) ? -1 : 0;
        };
        if (!d0_done && half(0, 0, grp(kGroupAcc0))) return -1;
        const bool meas = nout > 0 && (istep % nout) == 0 && row < max_rows;
        const bool chain = fuse2 && !meas && t + 1 < nsw;
        FA.istep = istep; FA.measure = meas ? 1 : 0; FA.row = row; FA.sweep_index = t;
        FA.close_prev = prev_open ? 1 : 0;
        if (fuse1) {
            set_u(0, 1);
            if (cfuse) {
                // colour 1's flips of this sweep go to the bitmap, which C0
                // of this sweep reads; its open trials to the queue
                A.flipbits = L->flipbits;
                A.held = L->held;
            }
            if (tbegin(L, KIND_FUSED, &ti)) return -1;
            launch_phase_fused(p->arith, rngk, L->bc, G, L->F, 1, A, grp(kGroupAcc1), grp(flip_group(istep, 1)),
                               L->kcfg[0][p->arith].tile, st);
            if (launch_check() || tend(L, ti)) return -1;
            if (cfuse) {
                // the held trials, decided on colour 0's chirality outcomes of
                // this sweep, before any colour-0 launch
                PhaseArgs A0 = A, A1 = A;
                if (p->rng != RNG_PHILOX) {
                    A0.u_acc = L->ustream + 4 * ns;
                    A1.u_acc = L->ustream + 5 * ns;
                }
                // in a chain, the finalize that closes sweep t's director
                // accepts shares the launch
                FA.close_this = 0;
                if (tbegin(L, KIND_HELD, &ti)) return -1;
                launch_held(rngk, L->bc, G, L->F, A0, A1, grp(flip_group(istep, 1)), L->held_total,
                            chain ? &FA : nullptr, st);
                if (launch_check() || tend(L, ti)) return -1;
            }
        } else if (half(0, 1, grp(kGroupAcc1))) {
            return -1;
        }
        if (chain) {
            // close sweep t's director accepts (cone step of sweep t + 1), then
            // C0 of sweep t with D0 of sweep t + 1
            FA.close_this = 0;
            if (!cfuse) {
                if (tbegin(L, KIND_FINAL, &ti)) return -1;
                launch_finalize(FA, st);
                if (launch_check() || tend(L, ti)) return -1;
            }
            PhaseArgs A2 = A;
            A2.istep = (uint32_t)(istep + 1);
            A2.istep_spin = (uint32_t)istep;
            // the director trials belong to sweep t + 1 and its temperature,
            // the chirality trials to sweep t
            A2.kbt = kt_of(istep + 1);
            A2.inv_kbt = 1.0 / A2.kbt;
            A2.inv_kbt_spin = A.inv_kbt;
            A2.kbt_spin = A.kbt;
            if (tbegin(L, KIND_FUSED_NEXT, &ti)) return -1;
            launch_phase_fused_next(p->arith, L->bc, G, L->F, A2, grp(kGroupAcc0), grp(flip_group(istep, 0)),
                                    L->kcfg[0][p->arith].tile, st);
            if (launch_check() || tend(L, ti)) return -1;
            d0_done = true;
            prev_open = true;
            continue;
        }
        if (half(1, 0, grp(flip_group(istep, 0)))) return -1;
        if (!fuse1 && half(1, 1, grp(flip_group(istep, 1)))) return -1;
        if (meas) {
            if (tbegin(L, KIND_MEASURE, &ti)) return -1;
            launch_measure(L->bc, G, L->F, p->g, p->KK, L->part, L->spart, st);
            if (launch_check() || tend(L, ti)) return -1;
        }
        FA.close_this = 1;
        if (tbegin(L, KIND_FINAL, &ti)) return -1;
        launch_finalize(FA, st);
        if (launch_check() || tend(L, ti)) return -1;
        if (meas) ++row;
        d0_done = false;
        prev_open = false;
    }
    if (rows && row > 0)
        HIPCHECK(hipMemcpyAsync(rows, L->rows_dev, (size_t)row * kRowCols * sizeof(double), hipMemcpyDeviceToHost, st));
    if (counts)
        HIPCHECK(hipMemcpyAsync(counts, L->counts_dev, (size_t)nsw * 2 * sizeof(long long), hipMemcpyDeviceToHost, st));
    if (rows || counts) HIPCHECK(hipStreamSynchronize(st));
    return row;
}

int chll_lattice_measure(chll_lattice *L, double g, double KK, double *row)
{
    if (!L || !row) return fail("chll_lattice_measure: null argument");
    if (ensure_rows(L, 1)) return -1;
    hipStream_t st = L->stream;
    // the accept/flip slots are zero between sweeps, so the row's counts read 0
    launch_measure(L->bc, L->G, L->F, g, KK, L->part, L->spart, st);
    if (launch_check()) return -1;
    FinArgs FA{};
    FA.counts = L->counts; FA.pend = L->pend; FA.part = L->part; FA.spart = L->spart;
    FA.nblk_meas = L->nblk_meas; FA.n3 = 2 * (uint64_t)L->G.nsub; FA.d = L->d_dev; FA.rows = L->rows_dev;
    FA.istep = 0; FA.measure = 1; FA.row = 0; FA.adapt = 0; FA.close_prev = 0; FA.close_this = 1;
    launch_finalize(FA, st);
    if (launch_check()) return -1;
    HIPCHECK(hipMemcpyAsync(row, L->rows_dev, kRowCols * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return 0;
}

// ---------------------------------------------------------------------------
// Drop-in entry points.

static void read_env_once()
{
    if (g_cfg_env_read) return;
    g_cfg_env_read = true;
    const char *q = getenv("CHLL_QUIET");
    if (q && *q && strcmp(q, "0") != 0) g_cfg.verbose = 0;
}

int chll_configure(int32_t rng, int32_t arith, int32_t bc, int32_t verbose)
{
    std::lock_guard<std::mutex> lk(g_mu);
    read_env_once();
    if (rng != RNG_PHILOX && rng != RNG_STREAM && rng != RNG_BUFFER) return fail("unknown rng mode %d", rng);
    if (arith != ARITH_COMPAT && arith != ARITH_FP64) return fail("unknown arithmetic mode %d", arith);
    if (bc != BC_OPEN && bc != BC_PERIODIC) return fail("unknown boundary condition %d", bc);
    g_cfg.rng = rng; g_cfg.arith = arith; g_cfg.bc = bc; g_cfg.verbose = verbose;
    return 0;
}

int chll_set_seed(uint64_t seed, uint32_t run_index)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_cfg.seed = seed;
    g_cfg.run_index = run_index;
    return 0;
}

int chll_set_stream(uint64_t seed, uint64_t pos)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_cfg.stream_seed = seed;
    g_cfg.stream_pos = pos;
    return 0;
}

uint64_t chll_get_stream_pos(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_cfg.stream_pos;
}

int chll_set_uniform_buffer(const double *u, int64_t n, uint64_t pos)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_cfg.buffer = u;
    g_cfg.buffer_len = u ? n : 0;
    g_cfg.stream_pos = pos;
    return 0;
}

int32_t chll_last_rows(double *out, int32_t rows)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!out) return fail("null output");
    const int32_t have = (int32_t)(g_cfg.last_rows.size() / kRowCols);
    const int32_t n = rows < have ? rows : have;
    if (n > 0) memcpy(out, g_cfg.last_rows.data(), (size_t)n * kRowCols * sizeof(double));
    return n;
}

void init(double *nx, double *ny, double *nz, int32_t *s, int32_t *dope, int32_t *sl1, int32_t *sl2,
          const int32_t *ni, const int32_t *nj, const int32_t *nk, const int32_t *nsub)
{
    (void)nx; (void)ny; (void)nz; (void)s; (void)dope;
    g_err.clear();
    if (!ni || !nj || !nk || !nsub || !sl1 || !sl2) { fail("init: null argument"); fprintf(stderr, "chll init: %s\n", g_err.c_str()); return; }
    const int64_t n3 = (int64_t)*ni * *nj * *nk;
    if (*ni < 1 || *nj < 1 || *nk < 1 || (n3 & 1) || *nsub != n3 / 2) {
        fail("init: need positive dims with N even and nsub = N/2 (ni,nj,nk,nsub = %d,%d,%d,%d)", *ni, *nj, *nk, *nsub);
        fprintf(stderr, "chll init: %s\n", g_err.c_str());
        return;
    }
    // chll/chll_mod.f90:23-36: Fortran order; 1-based i+j+k odd -> sl1, even -> sl2.
    int64_t a = 0, b = 0;
    for (int32_t k = 1; k <= *nk; ++k)
        for (int32_t j = 1; j <= *nj; ++j) {
            const int32_t base = (j - 1) * *ni + (k - 1) * (*ni * *nj);
            for (int32_t i = 1; i <= *ni; ++i) {
                if ((i + j + k) & 1) sl1[a++] = base + i;
                else sl2[b++] = base + i;
            }
        }
}

// Device lattices for run(): a call takes an idle one of its shape on the
// current device, or creates one, and gives it back on return, so calls from
// several host threads each sweep their own lattice.  One idle lattice per
// device is kept for the next call (the reference allocates on every call).
namespace {
std::mutex g_pool_mu;
std::vector<chll_lattice *> g_pool;
std::mutex g_print_mu;

chll_lattice *pool_take(int32_t ni, int32_t nj, int32_t nk, int bc)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        for (size_t i = 0; i < g_pool.size(); ++i) {
            chll_lattice *L = g_pool[i];
            if (L->device == dev && L->G.ni == (uint32_t)ni && L->G.nj == (uint32_t)nj &&
                L->G.nk == (uint32_t)nk && L->bc == bc) {
                g_pool.erase(g_pool.begin() + (long)i);
                return L;
            }
        }
    }
    chll_lattice *L = nullptr;
    if (chll_lattice_create(ni, nj, nk, bc, &L)) return nullptr;
    return L;
}

void pool_give(chll_lattice *L)
{
    std::vector<chll_lattice *> drop;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        for (size_t i = 0; i < g_pool.size();) {
            if (g_pool[i]->device == L->device) { drop.push_back(g_pool[i]); g_pool.erase(g_pool.begin() + (long)i); }
            else ++i;
        }
        g_pool.push_back(L);
    }
    for (chll_lattice *D : drop) chll_lattice_destroy(D);
}
}  // namespace

static void print_fixed(double v)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%12.6f", v);
    if (strlen(buf) > 12) printf(" ************");  // Fortran F12.6 overflow
    else printf(" %s", buf);
}

void run(const int32_t *nstart, const int32_t *nstop, const int32_t *nout,
         double *nx, double *ny, double *nz, int32_t *s, int32_t *dope, int32_t *sl1, int32_t *sl2,
         double *g, double *KK, double *d, double *kbt,
         const int32_t *ni, const int32_t *nj, const int32_t *nk, int32_t *nsub,
         int32_t *msteps, double *m)
{
    (void)dope; (void)sl1; (void)sl2;
    g_err.clear();
    auto bail = [](void) { fprintf(stderr, "chll run: %s\n", g_err.c_str()); };
    if (!nstart || !nstop || !nout || !nx || !ny || !nz || !s || !g || !KK || !d || !kbt || !ni || !nj || !nk ||
        !nsub) {
        fail("run: null argument");
        return bail();
    }
    if (*nout <= 0) { fail("run: nout must be positive (got %d)", *nout); return bail(); }
    const int32_t rows = (1 + *nstop - *nstart) / *nout;
    if (rows > 0 && (!msteps || !m)) { fail("run: null msteps/m"); return bail(); }
    // the options of this call, and its share of the random numbers: run
    // counters and stream positions are handed out under the lock, so
    // concurrent calls draw disjoint numbers
    RunConfig cfg;
    chll_params p{};
    {
        std::lock_guard<std::mutex> lk(g_mu);
        read_env_once();
        cfg.rng = g_cfg.rng; cfg.arith = g_cfg.arith; cfg.bc = g_cfg.bc; cfg.verbose = g_cfg.verbose;
        if (check_geometry(*ni, *nj, *nk, cfg.bc)) return bail();
        if (cfg.rng == RNG_BUFFER && !g_cfg.buffer) { fail("run: rng=buffer without chll_set_uniform_buffer"); return bail(); }
        const int64_t nsw = (int64_t)*nstop - *nstart + 1;
        const uint64_t draws = nsw > 0 ? (uint64_t)nsw * 10u * (((uint64_t)*ni * *nj * *nk) / 2) : 0;
        p.rng = cfg.rng; p.arith = cfg.arith;
        if (cfg.rng == RNG_PHILOX) { p.seed = g_cfg.seed; p.run_index = g_cfg.run_index++; }
        else {
            p.seed = g_cfg.stream_seed; p.stream_pos = g_cfg.stream_pos;
            if (cfg.rng == RNG_BUFFER && (int64_t)(p.stream_pos + draws) > g_cfg.buffer_len) {
                fail("uniform buffer too short: %lld sweeps need %llu draws from position %llu, buffer holds %lld",
                     (long long)nsw, (unsigned long long)draws, (unsigned long long)p.stream_pos,
                     (long long)g_cfg.buffer_len);
                return bail();
            }
            g_cfg.stream_pos += draws;
        }
        p.buffer = g_cfg.buffer; p.buffer_len = g_cfg.buffer_len;
    }
    chll_lattice *L = pool_take(*ni, *nj, *nk, cfg.bc);
    if (!L) return bail();
    *nsub = (int32_t)(((int64_t)*ni * *nj * *nk) / 2);
    p.g = *g; p.KK = *KK; p.kbt = *kbt; p.adapt = 1;
    std::vector<double> rbuf((size_t)(rows > 0 ? rows : 1) * kRowCols, 0.0);
    int got = -1;
    if (!chll_lattice_upload(L, nx, ny, nz, s) && !chll_lattice_set_step(L, *d)) {
        got = chll_lattice_sweeps(L, &p, *nstart, *nstop, *nout, rbuf.data(), rows > 0 ? rows : 0, nullptr);
        if (got >= 0 && (chll_lattice_download(L, nx, ny, nz, s) || chll_lattice_get_step(L, d))) got = -1;
    }
    pool_give(L);
    if (got < 0) return bail();
    for (int r = 0; r < got; ++r) {
        const double *row = &rbuf[(size_t)r * kRowCols];
        msteps[r] = (int32_t)row[0];
        for (int c = 0; c < 5; ++c) m[r + (size_t)c * rows] = row[1 + c];
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cfg.last_rows.assign(rbuf.begin(), rbuf.begin() + (size_t)got * kRowCols);
    }
    if (cfg.verbose) {
        // chll/chll_mod.f90:60 header and :98 rows, format (I8, 5(1X, F12.6))
        std::lock_guard<std::mutex> lk(g_print_mu);
        if (*nstart == 1) printf("Steps       Energy      E. excess      Paccept      Faccept      D\n");
        for (int r = 0; r < got; ++r) {
            const double *row = &rbuf[(size_t)r * kRowCols];
            printf("%8d", (int32_t)row[0]);
            for (int c = 1; c <= 5; ++c) print_fixed(row[c]);
            printf("\n");
        }
        fflush(stdout);
    }
}

}  // extern "C"
