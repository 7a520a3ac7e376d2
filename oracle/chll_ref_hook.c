/*
 * chll_ref_hook.c -- uniform-stream injection for the reference build.
 *
 * TEST INFRASTRUCTURE ONLY (linked into oracle/_ref/libchll_ref_stream.so,
 * never into the product).
 *
 * oracle/Makefile compiles the reference's own chll/chll_mod.f90 with the
 * preprocessor renaming the intrinsic calls random_number -> chll_hook_rn and
 * random_seed -> chll_hook_seed (chll/chll_mod.f90:58, 66-80).  Those become
 * implicit-interface external calls, so the hook receives the base address of
 * rand1 (nsub elements) or rand2 (2*nsub elements, column-major).  The
 * reference makes exactly six calls per sweep, of sizes nsub, 2nsub, nsub,
 * 2nsub, 2nsub, 2nsub, after one random_seed() per run, so the hook tracks the
 * call number within the sweep to know each array's length.
 *
 * The values come from the oracle's documented stream u_n
 * (oracle_stream_uniform) or from a caller-supplied buffer; either way the
 * position advances by 10*nsub per sweep, exactly as oracle_run consumes it.
 */
#include <stdint.h>
#include <stddef.h>

double oracle_stream_uniform(uint64_t seed, uint64_t n);

static int64_t g_nsub = 0;
static uint64_t g_seed = 0;
static uint64_t g_pos = 0;
static const double *g_buf = NULL;
static int64_t g_buf_len = 0;
static int g_call = 0;
static int g_overflow = 0;

/* Use the counter-based stream (seed) starting at draw index pos. */
void chll_ref_set_stream(uint64_t seed, uint64_t pos, int64_t nsub)
{
    g_seed = seed; g_pos = pos; g_nsub = nsub; g_buf = NULL; g_buf_len = 0; g_overflow = 0; g_call = 0;
}

/* Use an explicit buffer of uniforms starting at index pos. */
void chll_ref_set_buffer(const double *buf, int64_t len, uint64_t pos, int64_t nsub)
{
    g_buf = buf; g_buf_len = len; g_pos = pos; g_nsub = nsub; g_overflow = 0; g_call = 0;
}

uint64_t chll_ref_get_pos(void) { return g_pos; }
int chll_ref_overflow(void) { return g_overflow; }

void chll_hook_seed_(void) { g_call = 0; }

void chll_hook_rn_(double *a)
{
    static const int factor[6] = {1, 2, 1, 2, 2, 2};
    int64_t n = factor[g_call % 6] * g_nsub;
    for (int64_t t = 0; t < n; ++t) {
        uint64_t q = g_pos + (uint64_t)t;
        if (g_buf) {
            if ((int64_t)q < g_buf_len) a[t] = g_buf[q];
            else { a[t] = 2.0; g_overflow = 1; }
        } else {
            a[t] = oracle_stream_uniform(g_seed, q);
        }
    }
    g_pos += (uint64_t)n;
    ++g_call;
}
