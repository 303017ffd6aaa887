/*
 * ref_rng.c -- the uniform stream injected into the reference Fortran.
 *
 * TEST INFRASTRUCTURE ONLY (linked into oracle/_ref binaries).
 *
 * The reference draws its uniforms with `call random_number(x)` three times
 * per step, in the order time, proc, site
 * (kmos/fortran_src/proclist_generic_subroutines.mpy:22-25, 63-65, 98-100).
 * The build recipe compiles the checked-in generated Fortran with
 * -Drandom_number=kmcg_ref_uniform, so each of those calls lands here and
 * returns the same Philox4x32-10 uniforms as the engine and the oracle:
 * draw d belongs to step d/3, slot d%3; slots 0 and 1 are the two 64-bit
 * halves of block (step, 0), slot 2 the low half of block (step, 1).
 *
 * The timing variant without the per-step clock read is compiled with
 * -DCPU_TIME=kmcg_ref_noclock: update_clocks' `call CPU_TIME(runtime)`
 * (kmos/fortran_src/base.mpy:1144) then lands in kmcg_ref_noclock_, which
 * returns 0 (walltime is no result of the step loop).
 */
#include <stdint.h>

static uint32_t g_seed, g_stream;
static uint64_t g_draw;

static void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* called from the driver through bind(C) */
void kmcg_ref_rng_set(int32_t seed, int32_t stream)
{
    g_seed = (uint32_t)seed;
    g_stream = (uint32_t)stream;
    g_draw = 0;
}

/* continue the stream at the first draw of step `step` */
void kmcg_ref_rng_seek(int64_t step) { g_draw = 3 * (uint64_t)step; }

int64_t kmcg_ref_rng_draws(void) { return (int64_t)g_draw; }

/* implicit-interface Fortran call: `call kmcg_ref_uniform(x)` */
void kmcg_ref_uniform_(double *x)
{
    uint64_t step = g_draw / 3;
    int slot = (int)(g_draw % 3);
    uint32_t key[2] = {g_seed, g_stream};
    uint32_t ctr[4] = {(uint32_t)step, (uint32_t)(step >> 32), slot == 2 ? 1u : 0u, 0u};
    uint32_t o[4];
    philox(ctr, key, o);
    uint32_t lo = slot == 1 ? o[2] : o[0];
    uint32_t hi = slot == 1 ? o[3] : o[1];
    uint64_t v = ((uint64_t)hi << 32) | lo;
    *x = ((double)(v >> 12) + 0.5) * 0x1p-52;
    g_draw += 1;
}

/* implicit-interface Fortran call: `call kmcg_ref_noclock(runtime)` with a
 * default-kind real argument */
void kmcg_ref_noclock_(float *t) { *t = 0.0f; }
