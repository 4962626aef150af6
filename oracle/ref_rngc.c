/*
 * ref_rngc.c -- exports the reference's own C RNG restatement
 * (include/vsmc/rngc/{philox,threefry,u01}.h, compiled unchanged from
 * /root/reference/include) as plain functions, so the tests can check the
 * oracle's Philox4x32-10 / Threefry4x64-20 engines and U01 conversions
 * against reference code word for word. Built by oracle/Makefile into
 * oracle/_ref/libvsmc_rngc.so; TEST INFRASTRUCTURE ONLY.
 */
#include <vsmc/rngc/rngc.h>

/* n words of vsmc_philox4x32_rand after vsmc_philox4x32_init(seed)
 * (rngc/philox.h:246-281), skipping `skip` words first */
void ref_philox4x32_words(uint32_t seed, uint64_t skip, size_t n,
    uint32_t *out)
{
    vsmc_philox4x32 rng;
    vsmc_philox4x32_init(&rng, seed);
    for (uint64_t i = 0; i < skip; ++i)
        (void)vsmc_philox4x32_rand(&rng);
    for (size_t i = 0; i < n; ++i)
        out[i] = vsmc_philox4x32_rand(&rng);
}

/* one block: vsmc_philox4x32_gen (rngc/philox.h:206-232) */
void ref_philox4x32_block(const uint32_t ctr[4], const uint32_t key[2],
    uint32_t out[4])
{
    vsmc_philox4x32_ctr_t c, s;
    vsmc_philox4x32_key_t k;
    for (int i = 0; i < 4; ++i)
        c.v[i] = ctr[i];
    k.v[0] = key[0];
    k.v[1] = key[1];
    vsmc_philox4x32_gen(&c, &k, &s);
    for (int i = 0; i < 4; ++i)
        out[i] = s.v[i];
}

/* rngc/threefry.h:497-561 */
void ref_threefry4x64_words(uint64_t seed, uint64_t skip, size_t n,
    uint64_t *out)
{
    vsmc_threefry4x64 rng;
    vsmc_threefry4x64_init(&rng, seed);
    for (uint64_t i = 0; i < skip; ++i)
        (void)vsmc_threefry4x64_rand(&rng);
    for (size_t i = 0; i < n; ++i)
        out[i] = vsmc_threefry4x64_rand(&rng);
}

/* rngc/threefry.h:420-453 */
void ref_threefry4x64_block(const uint64_t ctr[4], const uint64_t key[4],
    uint64_t out[4])
{
    vsmc_threefry4x64_ctr_t c, s;
    vsmc_threefry4x64_key_t k;
    for (int i = 0; i < 4; ++i) {
        c.v[i] = ctr[i];
        k.v[i] = key[i];
    }
    vsmc_threefry4x64_gen(&c, &k, &s);
    for (int i = 0; i < 4; ++i)
        out[i] = s.v[i];
}

/* rngc/u01.h:214-217, :243-246 */
void ref_u01_co_u32d(size_t n, const uint32_t *u, double *out)
{
    for (size_t i = 0; i < n; ++i)
        out[i] = vsmc_u01_co_u32d(u[i]);
}

void ref_u01_co_u64d(size_t n, const uint64_t *u, double *out)
{
    for (size_t i = 0; i < n; ++i)
        out[i] = vsmc_u01_co_u64d(u[i]);
}

void ref_u01_co_u32f(size_t n, const uint32_t *u, float *out)
{
    for (size_t i = 0; i < n; ++i)
        out[i] = vsmc_u01_co_u32f(u[i]);
}
