/*
 * chll_mod.h -- C ABI of libchll_mod.so, the MI355X (gfx950) implementation
 * of the chiral Lebwohl-Lasher Metropolis sweep of mdeutsc2/ChLLModel.
 *
 * The two drop-in entry points keep the reference's symbol names, argument
 * order, by-reference scalars and Fortran-ordered array layouts, so the
 * reference's Python binding (chll/chll.py:6-36) loads this library unchanged.
 * Everything prefixed chll_ is an extension with no counterpart in the
 * reference.  Extension functions return 0 on success and -1 on error, with
 * the message available from chll_last_error().
 */
#ifndef CHLL_MOD_H
#define CHLL_MOD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Drop-in entry points                                                     */
/* ------------------------------------------------------------------------ */

/* Replaces `subroutine init(nx,ny,nz,s,dope,sl1,sl2,ni,nj,nk,nsub)
 * bind(c, name="init")`, chll/chll_mod.f90:9-38 (Python argtypes
 * chll/chll.py:9-18).  Fills sl1[0..nsub) / sl2[0..nsub) with the 1-based
 * Fortran linear indices of the sites whose 1-based i+j+k is odd / even, in
 * Fortran order.  The other arrays are not touched.  Requires N = ni*nj*nk
 * even and nsub = N/2 (the reference overflows sl1 otherwise); on a bad size
 * nothing is written and chll_last_error() says why. */
void init(double *nx, double *ny, double *nz, int32_t *s, int32_t *dope, int32_t *sl1, int32_t *sl2,
          const int32_t *ni, const int32_t *nj, const int32_t *nk, const int32_t *nsub);

/* Replaces `subroutine run(nstart,nstop,nout,nx,ny,nz,s,dope,sl1,sl2,g,KK,d,
 * kbt,ni,nj,nk,nsub,msteps,m) bind(c, name="run")`, chll/chll_mod.f90:40-115
 * (Python argtypes chll/chll.py:20-36).  Runs sweeps istep = nstart..nstop on
 * the GPU: director(sl1), director(sl2), chirality(sl1), chirality(sl2) per
 * sweep; adapts d while istep < 10000; every nout sweeps records
 * msteps[r] = istep and the column-major row m[r + c*rows] =
 * (E/site, eta, paccept, faccept, d), rows = (1+nstop-nstart)/nout, and prints
 * the reference's table to stdout (silenced by chll_configure(..., verbose=0)
 * or CHLL_QUIET=1).  nx, ny, nz, s are updated in place, *d receives the
 * adapted step and *nsub = N/2.  Arrays are Fortran-ordered (ni,nj,nk),
 * any shape with N even (periodic boundaries: ni, nj, nk even).  No device
 * state survives the call.  Boundary condition, arithmetic and random source
 * are those set by chll_configure (default: open boundaries,
 * reference-compatible arithmetic, Philox).  Re-entrant: calls from several
 * host threads each sweep their own device lattice. */
void run(const int32_t *nstart, const int32_t *nstop, const int32_t *nout,
         double *nx, double *ny, double *nz, int32_t *s, int32_t *dope, int32_t *sl1, int32_t *sl2,
         double *g, double *KK, double *d, double *kbt,
         const int32_t *ni, const int32_t *nj, const int32_t *nk, int32_t *nsub,
         int32_t *msteps, double *m);

/* ------------------------------------------------------------------------ */
/* Extensions                                                               */
/* ------------------------------------------------------------------------ */

#define CHLL_RNG_PHILOX 0 /* site-keyed Philox4x32-10 counters on the device */
#define CHLL_RNG_STREAM 1 /* documented uniform stream u_n in the reference's draw schedule */
#define CHLL_RNG_BUFFER 2 /* caller's host array of uniforms in the reference's draw schedule */

#define CHLL_ARITH_COMPAT 0 /* reference arithmetic: binary32-rounded site energies */
#define CHLL_ARITH_FP64 1   /* binary64 energies, fused multiply-adds */

#define CHLL_BC_OPEN 0     /* free boundaries, as the reference's sweep */
#define CHLL_BC_PERIODIC 1 /* periodic in i, j, k (ni, nj, nk even) */

/* Measurement row columns (chll_lattice_sweeps, chll_last_rows). */
#define CHLL_ROW_COLS 16
/* 0 istep | 1 E: reference etot | 2 eta | 3 paccept | 4 faccept | 5 d |
 * 6 E_true/N: Hamiltonian of the sweep under its boundary condition |
 * 7 exact mean spin | 8..13 <nx nx>, <ny ny>, <nz nz>, <nx ny>, <nx nz>, <ny nz> |
 * 14 naccept | 15 nflip */

/* Message of the last failed call (empty string when none). */
const char *chll_last_error(void);

/* Options of subsequent run() calls.  verbose: 1 prints the reference's table. */
int chll_configure(int32_t rng, int32_t arith, int32_t bc, int32_t verbose);

/* Philox key and run counter for run(); run_index advances by one per run()
 * call so consecutive calls draw fresh numbers. */
int chll_set_seed(uint64_t seed, uint32_t run_index);

/* CHLL_RNG_STREAM: stream seed and next draw index (advances 10*nsub per sweep). */
int chll_set_stream(uint64_t seed, uint64_t pos);
uint64_t chll_get_stream_pos(void);

/* CHLL_RNG_BUFFER: host array of n uniforms, next draw index pos.  The array
 * must stay valid while run() consumes it. */
int chll_set_uniform_buffer(const double *u, int64_t n, uint64_t pos);

/* Copy up to `rows` full measurement rows (row-major, CHLL_ROW_COLS each) of
 * the last run() call; returns the number copied or -1. */
int32_t chll_last_rows(double *out, int32_t rows);

/* Device-resident lattice -------------------------------------------------- */

typedef struct chll_params {
    double g, KK, kbt;
    int32_t rng, arith;
    uint64_t seed;       /* Philox key, or the stream seed */
    uint64_t stream_pos; /* STREAM / BUFFER: next draw index, advanced on return */
    const double *buffer;/* BUFFER: host uniforms */
    int64_t buffer_len;
    uint32_t run_index, replica;
    int32_t adapt;       /* 1: reference step adaptation while istep < 10000 */
    int32_t pad;
    double kbt_end;      /* 0: every sweep at kbt.  > 0: a linear anneal, sweep istep at
                            kbt + (kbt_end - kbt) * (double)(istep - nstart) / (double)(nstop - nstart)
                            (binary64, in this order), sweep nstop at kbt_end exactly (kbt
                            itself when nstop == nstart).  No
                            counterpart in the reference, which holds kbt for a whole run
                            (chll/chll_mod.f90:40, scripts/pdanneal.py:21-37 steps it per run) */
} chll_params;

typedef struct chll_lattice chll_lattice;

int chll_lattice_create(int32_t ni, int32_t nj, int32_t nk, int32_t bc, chll_lattice **out);
int chll_lattice_destroy(chll_lattice *lat);
/* Launch on the caller's HIP stream (NULL restores the lattice's own stream). */
int chll_lattice_set_stream(chll_lattice *lat, void *hip_stream);
/* Host Fortran-ordered fields <-> device colour-split layout. */
int chll_lattice_upload(chll_lattice *lat, const double *nx, const double *ny, const double *nz, const int32_t *s);
int chll_lattice_download(chll_lattice *lat, double *nx, double *ny, double *nz, int32_t *s);
/* Initial states generated on the device, replacing the reference's Python
 * loops (scripts/pdanneal.py:5-16 init_random, scripts/hero480.py:18-21
 * init_aligned2, chll/chll.py:84-98 init_rand_s and its default n = (1,0,0)).
 * Site with Fortran index l (0-based) draws Philox4x32-10 blocks keyed by
 * seed with counters (l, replica, 0x494E4954, 0) and (..., 1): isotropic
 * directors costh = 2 u53(x0,y0) - 1, phi = 2 pi u53(z0,w0); spin -1 when
 * u53(x1,y1) < p_flip, else +1.  The state depends only on (seed, replica,
 * site), not on the kernel or layout. */
#define CHLL_INIT_ISOTROPIC 0
#define CHLL_INIT_ALIGNED_X 1  /* n = (1, 0, 0) */
#define CHLL_INIT_ALIGNED_XZ 2 /* n = (1, 0, 1) / sqrt(2) */
int chll_lattice_init_state(chll_lattice *lat, int32_t kind, uint64_t seed, uint32_t replica, double p_flip);
int chll_lattice_set_step(chll_lattice *lat, double d);
int chll_lattice_get_step(chll_lattice *lat, double *d);
/* Sweeps istep = nstart..nstop.  A row is measured when nout > 0 and
 * istep % nout == 0, up to max_rows rows.  When rows or counts (2 per sweep:
 * naccept, nflip) is non-NULL the call waits and copies them; otherwise it
 * only enqueues work. */
int chll_lattice_sweeps(chll_lattice *lat, chll_params *p, int32_t nstart, int32_t nstop, int32_t nout,
                        double *rows, int32_t max_rows, int64_t *counts);
/* Measure one row of the current state without sweeping (row-major, CHLL_ROW_COLS). */
int chll_lattice_measure(chll_lattice *lat, double g, double KK, double *row);
int chll_lattice_sync(chll_lattice *lat);
/* Half-sweep kernel for one trial kind (0 director, 1 chirality) and
 * arithmetic: variant 0 = one site per thread, tile bit 0 = each of the 8 XCDs
 * takes one band of rows of every plane (else blocks in lattice order, also
 * when a plane does not split into 8 bands of whole blocks), tile bit 1 = two
 * neighbouring sites per thread read with 16-byte pair loads (when ni % 4 == 0
 * and a field spans less than 4 GiB; else one site per thread), tile bit 2
 * (director kind, FP64 fused sweeps with site pairs) = the fused colour-0
 * launch walks its blocks in reverse order;
 * variant 1 = tiled through LDS: tile
 * 0 = 64x4, 1 = 32x8, 2 = 16x16 packed sites x rows marching along k in nslab
 * slabs, 3..5 = the same tile shapes one plane per block (nslab unused).
 * Results do not depend on the choice. */
int chll_lattice_set_kernel(chll_lattice *lat, int32_t kind, int32_t arith, int32_t variant, int32_t tile,
                            int32_t nslab);
int chll_lattice_get_kernel(chll_lattice *lat, int32_t kind, int32_t arith, int32_t *variant, int32_t *tile,
                            int32_t *nslab);
/* FP64 sweeps (on by default): each colour-1 site's chirality trial runs in
 * the colour-1 director launch, right after that site's director trial, when
 * the director kernel is the one-site kernel (variant 0); with Philox draws
 * each colour-0 chirality trial also shares a launch with the same site's
 * director trial of the next sweep (not across a measurement sweep or the end
 * of the call).  FP64 chirality energy changes do not read neighbour spins,
 * so the lattice, counts and rows are the same as with four separate
 * half-sweeps.  No counterpart in the reference (chll/chll_mod.f90:70-81 runs
 * the four half-sweeps in turn). */
int chll_lattice_set_fusion(chll_lattice *lat, int32_t on);
int chll_lattice_get_fusion(chll_lattice *lat, int32_t *on);
/* Reference arithmetic (CHLL_ARITH_COMPAT): each trial is first decided from
 * its binary64 energy difference, which lies within scale x M of the
 * difference of the reference's binary32-rounded site energies (M bounds the
 * two energies' terms; scale 2^-21, the default, covers the roundings of
 * chll/chll_mod.f90:363-412); only trials whose outcome that margin leaves
 * open evaluate the rounded energies.  The decisions, and so the lattice,
 * counts and rows, are the reference arithmetic's at every scale.  scale = 0
 * rounds both energies of every trial; scales in (2^-21, 1] send more trials
 * to them; others are rejected.  No counterpart in the reference. */
int chll_lattice_set_screen(chll_lattice *lat, double scale);
/* Reference-arithmetic sweeps run colour 1's chirality trials inside its
 * director launch where the screen is on, kT > 0 and both kernels take site
 * pairs (fusion on): a trial whose outcome the screen leaves open is held and
 * decided after colour 0's chirality outcomes are known, as the reference's
 * order (chll/chll_mod.f90:70-81) has it.  Writes the number of trials held
 * on this lattice since its creation.  No counterpart in the reference. */
int chll_lattice_held_trials(chll_lattice *lat, int64_t *n);
/* Per-kernel HIP-event timing: kinds 0 director, 1 chirality, 2 measure,
 * 3 finalize, 4 uniform fill, 5 fused colour-1 director + chirality, 6 fused
 * colour-0 chirality + next sweep's colour-0 director, 7 held chirality
 * trials (decision and flips). */
int chll_lattice_set_timing(chll_lattice *lat, int32_t on);
int chll_lattice_timing(chll_lattice *lat, double *ms, int64_t *launches, int32_t nkinds);

/* Measurement helper, no counterpart in the reference: the rate a plain
 * streaming kernel reaches for one mix of loads and stores, the ceiling the
 * sweep kernels' algorithmic rate is compared with.  Allocates two device
 * arrays of `bytes` each on the current device (256 MiB to 16 GiB; 2 GiB puts
 * them well past the 256 MB memory-side cache), zero-fills them, runs 3
 * untimed and then `reps` (1..1000) timed passes on a stream of its own, frees
 * the arrays and writes the best and median rates in GB/s (10^9 B/s) of bytes
 * moved per pass.  Waits for the GPU.  Patterns: */
#define CHLL_BW_COPY 0         /* read a, write b (2 x bytes per pass) */
#define CHLL_BW_READ2_WRITE1 1 /* read a and b, write a in place (3 x bytes): the
                                  sweep's mix, its own colour read and rewritten
                                  and the other colour read */
#define CHLL_BW_READ 2         /* read a (1 x bytes) */
#define CHLL_BW_WRITE 3        /* write b (1 x bytes) */
int chll_bandwidth_probe(int32_t pattern, int64_t bytes, int32_t reps, double *best_gbps, double *median_gbps);

/* Batched replicas with parallel tempering ------------------------------- */
/* No counterpart in the reference, which scans (kT, K, g) grids by running one
 * ChLLSim after another (e.g. scripts/phase_diagram_large1.py:30-37). */

#define CHLL_ENS_COLS 16
/* ensemble row columns: 0 istep | 1 kT | 2 E_true/N | 3 mean spin | 4..9 <n_a n_b>
 * (xx, yy, zz, xy, xz, yz) | 10 naccept/N | 11 nflip/N | 12 d | 13 temperature
 * slot | 14 reference etot (binary32) | 15 reserved */

typedef struct chll_replica_desc {
    int32_t ni, nj, nk;
    int32_t ladder; /* replicas of one ladder: same ni, nj, nk, g, KK; listed
                       contiguously in ascending kbt */
    double g, KK, kbt, d0;
} chll_replica_desc;

typedef struct chll_ensemble chll_ensemble;

const char *chll_ensemble_last_error(void);
int chll_ensemble_create(const chll_replica_desc *reps, int32_t nrep, int32_t bc, int32_t arith, chll_ensemble **out);
int chll_ensemble_destroy(chll_ensemble *e);
int chll_ensemble_set_stream(chll_ensemble *e, void *hip_stream);
/* threads per block of the batched kernels: 64, 128, 256 or 512 */
int chll_ensemble_set_block(chll_ensemble *e, int32_t block_size);
/* on = 1 (default) runs the fused launches of chll_lattice_set_fusion inside
 * each block of K sweeps, two sites per thread when every replica has
 * ni % 4 == 0 (the single lattice's site pairs); 2 the same with one site per
 * thread; 0 four half-sweeps.  Reference arithmetic fuses only with 1 on site
 * pairs, through held chirality trials, and otherwise runs four half-sweeps.
 * Results are unchanged. */
int chll_ensemble_set_fusion(chll_ensemble *e, int32_t on);
/* Replicas are independent between swaps, so a block of K sweeps can run chunk
 * by chunk: all K sweeps of consecutive replicas (largest first) holding at
 * most `bytes` of lattice state (25 B per site), then the next chunk, so that
 * a chunk stays in the memory-side cache for its K sweeps.  0 (default): one
 * chunk.  Results are unchanged.  chll_ensemble_chunks: the number of chunks. */
int chll_ensemble_set_chunk(chll_ensemble *e, int64_t bytes);
int32_t chll_ensemble_chunks(chll_ensemble *e);
/* With n = 2 or 3 the chunks take turns on n streams forked from and joined
 * back to the ensemble's stream (n branches of the captured graph), so one
 * chunk's short launches run beside another's sweeps; 1 (default): one after
 * another.  Results are unchanged.  No counterpart in the reference (its scans
 * run one ChLLSim after another, e.g. scripts/phase_diagram_large1.py). */
int chll_ensemble_set_branches(chll_ensemble *e, int32_t n);
int32_t chll_ensemble_blocks(chll_ensemble *e);
int chll_ensemble_upload(chll_ensemble *e, int32_t rep, const double *nx, const double *ny, const double *nz,
                         const int32_t *s);
int chll_ensemble_download(chll_ensemble *e, int32_t rep, double *nx, double *ny, double *nz, int32_t *s);
/* chll_lattice_init_state for every replica, replica index r as its counter word. */
int chll_ensemble_init_state(chll_ensemble *e, int32_t kind, uint64_t seed, double p_flip);
/* nblocks blocks of K sweeps; each block ends with a measurement row per
 * replica and a parallel-tempering swap attempt in every ladder.  rows:
 * [min(nblocks, max_rows)][nrep][CHLL_ENS_COLS]; NULL rows only enqueues.
 * use_graph: capture one block in a hipGraph and replay it. */
int chll_ensemble_run(chll_ensemble *e, uint64_t seed, uint32_t run_index, int32_t nblocks, int32_t K, int32_t adapt,
                      int32_t use_graph, double *rows, int32_t max_rows);
int chll_ensemble_sync(chll_ensemble *e);
/* replica -> slot, slot kT and step, swap attempts/acceptances per slot pair
 * (t, t+1), next sweep index; any pointer may be NULL */
int chll_ensemble_state(chll_ensemble *e, int32_t *rep_slot, double *slot_kbt, double *slot_d, int64_t *swap_att,
                        int64_t *swap_acc, uint32_t *istep);

/* Lattice energy and gradient ------------------------------------------ */
/* No counterpart in the reference (its Hamiltonian, chll/chll_mod.f90:369-411,
 * is only evaluated inside the Metropolis test).  Device pointers, batch of B
 * lattices in plain Fortran order: director component c of lattice b at
 * n + (b*3 + c)*N, int8 spins at s + b*N, per-lattice eps, K, g.  Computes
 * E_b = sum_bonds [eps (1 - c^2) - K c q f] - g sum s (each bond once, open or
 * periodic boundaries) and, for any non-NULL output, dE/dn projected on the
 * tangent plane of each director (B x 3 x N), dE/dK, dE/dg, dE/deps (B each).
 * workspace: device memory of chll_energy_workspace_bytes(); stream: HIP
 * stream or NULL.  Enqueues only. */
const char *chll_energy_last_error(void);
size_t chll_energy_workspace_bytes(int32_t batch, int32_t ni, int32_t nj, int32_t nk);
int chll_energy_grad(const double *n, const int8_t *s, int32_t batch, int32_t ni, int32_t nj, int32_t nk, int32_t bc,
                     const double *K, const double *g, const double *eps, double *E, double *dEdn, double *dEdK,
                     double *dEdg, double *dEdeps, void *workspace, size_t workspace_bytes, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation): chll_energy_grad with its workspace taken from a static
 * device buffer of CHLL_ENERGY_SCRATCH_BYTES per GPU (held until the process
 * ends) instead of a caller argument.  Same kernels and results.  Fails, with
 * chll_energy_last_error() set and nothing launched, when
 * chll_energy_workspace_bytes(batch, ni, nj, nk) exceeds the buffer.
 * Allocates nothing.  All calls on one GPU share the buffer: calls that
 * overlap on one GPU (concurrent streams or host threads) race; calls in one
 * stream's order do not. */
#define CHLL_ENERGY_SCRATCH_BYTES ((size_t)8 << 20)
int chll_energy_grad_noalloc(const double *n, const int8_t *s, int32_t batch, int32_t ni, int32_t nj, int32_t nk,
                             int32_t bc, const double *K, const double *g, const double *eps, double *E, double *dEdn,
                             double *dEdK, double *dEdg, double *dEdeps, void *stream);

/* Analysis ------------------------------------------------------------- */
/* Device pointers, one lattice in plain Fortran order (i fastest), float64
 * director components, int8 spins.  Functions return 0 or -1 with
 * chll_analysis_last_error() set; the device work is enqueued on `stream`
 * (HIP stream or NULL) and the call does not wait for it.  The small host
 * tables each call copies go through a pinned staging buffer per host thread;
 * a call waits only for that buffer's previous copies.  The workspace is used
 * in stream order. */
const char *chll_analysis_last_error(void);

/* Replaces calc_n_corr and calc_s_corr (notebooks/paper_plots2.ipynb, cell 0;
 * numba loops).  For shells sh = 0 .. nshells-1 (1 <= nshells <= 127) writes
 * g_nem[sh] = mean of (n_r . n_{r+d})^2 and g_chi[sh] = mean of
 * s_r s_{r+d} - <s>^2 over the ordered pairs (r, r + d) inside the lattice
 * with floor(|d|) = sh and |d| < nshells; either output may be NULL (then its
 * inputs may be too).  Fails when a shell has no pairs (the reference divides
 * by zero there).  workspace: chll_corr_workspace_bytes(). */
size_t chll_corr_workspace_bytes(int32_t ni, int32_t nj, int32_t nk, int32_t nshells);
int chll_shell_corr(const double *nx, const double *ny, const double *nz, const int8_t *s, int32_t ni, int32_t nj,
                    int32_t nk, int32_t nshells, double *g_nem, double *g_chi, void *workspace,
                    size_t workspace_bytes, void *stream);
/* Host only: the pair counts per shell (the reference's second return value). */
int chll_corr_counts(int32_t ni, int32_t nj, int32_t nk, int32_t nshells, int64_t *counts);

/* Replaces get_sc / structure_factor
 * (notebooks/.ipynb_checkpoints/hero_stripe-checkpoint.py:10-29; numba loop).
 * planes: host array of 3 ints per plane (1 <= nplanes <= 16); kv: host array
 * of `bins` wave numbers.  Writes sf[p*bins + b] =
 * |sum_r s_r exp(-i kv[b] planes[p] . (r + 1))|^2 / N^2 (device, nplanes x bins),
 * r = (i, j, k) 0-based.  The planes' projections p . (r + 1) may span at most
 * 12288 distinct values in total per call.
 * workspace: chll_sf_workspace_bytes() (0 for invalid planes). */
size_t chll_sf_workspace_bytes(int32_t ni, int32_t nj, int32_t nk, const int32_t *planes, int32_t nplanes,
                               int32_t bins);
int chll_structure_factor(const int8_t *s, int32_t ni, int32_t nj, int32_t nk, const int32_t *planes,
                          int32_t nplanes, const double *kv, int32_t bins, double *sf, void *workspace,
                          size_t workspace_bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CHLL_MOD_H */
