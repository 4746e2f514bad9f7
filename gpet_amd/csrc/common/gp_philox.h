/*
 * gp_philox.h - Philox4x32-10 counter-based generator (Salmon, Moraes, Dror,
 * Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), shared by the
 * HIP kernels and the host oracle so both draw identical streams.
 *
 * Replaces the reference's per-thread cuRAND XORWOW states seeded from
 * srand(time(NULL)) (initialize.cu:246-275, gPET_kernals.cu:7-17).  Every
 * photon, decay and digitizer event owns a stream, so the numbers a history
 * consumes do not depend on launch shape, batching or regrouping.
 *
 * Counter layout: { block, aux, id_lo, id_hi }, key = 64-bit seed.  With
 * aux == 0 this matches rocRAND's philox4x32_10_engine seeded with `seed`,
 * subsequence `id` and offset 4*block.
 */
#ifndef GP_PHILOX_H
#define GP_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#define GP_HD __host__ __device__ __forceinline__
#else
#define GP_HD static inline
#endif

/* Three-way xor: one v_bitop3_b32 (truth table 0x96) in gfx950 code, where
 * the compiler would otherwise issue two v_xor_b32; plain C elsewhere. */
#if defined(__HIP_DEVICE_COMPILE__)
#define GP_XOR3(a, b, c) __builtin_amdgcn_bitop3_b32((a), (b), (c), 0x96)
#else
#define GP_XOR3(a, b, c) ((a) ^ (b) ^ (c))
#endif

#define GP_PHILOX_M0 0xD2511F53u
#define GP_PHILOX_M1 0xCD9E8D57u
#define GP_PHILOX_W0 0x9E3779B9u
#define GP_PHILOX_W1 0xBB67AE85u

typedef struct gp_u4 {
    uint32_t x, y, z, w;
} gp_u4;

GP_HD gp_u4 gp_philox4x32_10(gp_u4 c, uint32_t k0, uint32_t k1)
{
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)GP_PHILOX_M0 * (uint64_t)c.x;
        uint64_t p1 = (uint64_t)GP_PHILOX_M1 * (uint64_t)c.z;
        gp_u4 n;
        n.x = GP_XOR3((uint32_t)(p1 >> 32), c.y, k0);
        n.y = (uint32_t)p1;
        n.z = GP_XOR3((uint32_t)(p0 >> 32), c.w, k1);
        n.w = (uint32_t)p0;
        c = n;
        k0 += GP_PHILOX_W0;
        k1 += GP_PHILOX_W1;
    }
    return c;
}

/* Stream tags (aux low 4 bits). */
enum {
    GP_TAG_DECAY = 0,     /* source, decay time, positron, pair direction */
    GP_TAG_PHOTON = 1,    /* aux = 1 + photon index (0/1) -> 1, 2          */
    GP_TAG_DIGI = 4,      /* digitizer: aux = 4 | (siten << 4), id = parn  */
    GP_TAG_NOISE_T = 5,   /* background arrival gaps: aux = 5 | (job << 4), id = chunk */
    GP_TAG_NOISE_A = 6    /* their crystal and energy: aux = 6 | (job << 4), id = chunk,
                             draws 8 m .. 8 m + 7 for the chunk's m-th single  */
};

/* One stream: 32-bit draw counter plus the 4-word block it is reading. */
typedef struct gp_rng {
    uint32_t aux, id_lo, id_hi, k0, k1;
    uint32_t n;              /* draws consumed                             */
    uint32_t b0, b1, b2, b3; /* current block (valid when n % 4 != 0)      */
} gp_rng;

GP_HD void gp_rng_init(gp_rng *s, uint64_t seed, uint64_t id, uint32_t aux, uint32_t n)
{
    s->aux = aux;
    s->id_lo = (uint32_t)id;
    s->id_hi = (uint32_t)(id >> 32);
    s->k0 = (uint32_t)seed;
    s->k1 = (uint32_t)(seed >> 32);
    s->n = n;
    s->b0 = s->b1 = s->b2 = s->b3 = 0u;
    if (n & 3u) {
        gp_u4 c = {n >> 2, aux, s->id_lo, s->id_hi};
        gp_u4 r = gp_philox4x32_10(c, s->k0, s->k1);
        s->b0 = r.x; s->b1 = r.y; s->b2 = r.z; s->b3 = r.w;
    }
}

GP_HD uint32_t gp_rng_u32(gp_rng *s)
{
    uint32_t i = s->n & 3u;
    if (i == 0u) {
        gp_u4 c = {s->n >> 2, s->aux, s->id_lo, s->id_hi};
        gp_u4 r = gp_philox4x32_10(c, s->k0, s->k1);
        s->b0 = r.x; s->b1 = r.y; s->b2 = r.z; s->b3 = r.w;
    }
    s->n += 1u;
    return i == 0u ? s->b0 : (i == 1u ? s->b1 : (i == 2u ? s->b2 : s->b3));
}

/* Skip to the next 4-word block boundary.  Called at the start of every
 * sampling stage, so lanes whose earlier rejection loops ran for different
 * numbers of iterations refill their blocks at the same draw sites (on a
 * GPU a refill then runs once per wavefront instead of once per divergent
 * lane group). */
GP_HD void gp_rng_align(gp_rng *s) { s->n = (s->n + 3u) & ~3u; }

/* The whole next block (after aligning): one generator call for up to four
 * draws of a stage. */
GP_HD gp_u4 gp_rng_block(gp_rng *s)
{
    gp_rng_align(s);
    gp_u4 c = {s->n >> 2, s->aux, s->id_lo, s->id_hi};
    s->n += 4u;
    return gp_philox4x32_10(c, s->k0, s->k1);
}

/* Uniform in (0, 1]: 24 random bits, same range as curand_uniform. */
GP_HD float gp_u01(uint32_t x)
{
    return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f; /* 2^-24 */
}

GP_HD float gp_rng_uniform(gp_rng *s) { return gp_u01(gp_rng_u32(s)); }

/* Uniform in (0, 1] with 53 random bits (two draws). */
GP_HD double gp_rng_uniform_d(gp_rng *s)
{
    uint32_t a = gp_rng_u32(s) >> 5;   /* 27 bits */
    uint32_t b = gp_rng_u32(s) >> 6;   /* 26 bits */
    return ((double)a * 67108864.0 + (double)b + 1.0) * 1.1102230246251565e-16; /* 2^-53 */
}

#endif /* GP_PHILOX_H */
