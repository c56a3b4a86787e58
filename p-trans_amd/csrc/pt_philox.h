/*
 * pt_philox.h -- counter-based random numbers for the phonon trace.
 *
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies and a Weyl key schedule.
 *
 * Stream layout.  Every phonon path owns its own stream, so a path can be traced,
 * re-traced (adjoint) or replayed on any lane, in any order, on the CPU or the GPU:
 *   key     = (seed lo32, seed hi32)
 *   counter = (block lo32, block hi32, path lo32, path hi32)
 * Draw d of a path uses block d>>1; one block gives two doubles, words (0,1) for even
 * d and words (2,3) for odd d.  A double takes 53 random bits:
 *   u = ((w0 >> 5) * 2^26 + (w1 >> 6)) * 2^-53,   u in [0, 1 - 2^-53]
 * which is exactly representable, so the CPU and GPU see identical values.
 * The reference draws from the compiler's RANDOM_NUMBER, also on [0,1)
 * (src/RTMCphon.f90:1850-1851, 1866, 2171, 2197-2198, 2213, 2218, 2230, 2361, 2399, 2410).
 */
#ifndef PT_PHILOX_H
#define PT_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#define PT_HD __host__ __device__ static inline
#else
#define PT_HD static inline
#endif

#define PT_PHILOX_M0 0xD2511F53u
#define PT_PHILOX_M1 0xCD9E8D57u
#define PT_PHILOX_W0 0x9E3779B9u
#define PT_PHILOX_W1 0xBB67AE85u

PT_HD void pt_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                            uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    int r;
    for (r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PT_PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PT_PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += PT_PHILOX_W0;
        k1 += PT_PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

PT_HD double pt_u01_from_words(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* Both doubles of block `block` of path `path`. */
PT_HD void pt_philox_block(uint64_t seed, uint64_t path, uint64_t block, double u[2])
{
    uint32_t ctr[4], key[2], out[4];
    ctr[0] = (uint32_t)block;
    ctr[1] = (uint32_t)(block >> 32);
    ctr[2] = (uint32_t)path;
    ctr[3] = (uint32_t)(path >> 32);
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
    pt_philox4x32_10(ctr, key, out);
    u[0] = pt_u01_from_words(out[0], out[1]);
    u[1] = pt_u01_from_words(out[2], out[3]);
}

/* Draw number `draw` (0-based) of path `path`. */
PT_HD double pt_philox_draw(uint64_t seed, uint64_t path, uint64_t draw)
{
    double u[2];
    pt_philox_block(seed, path, draw >> 1, u);
    return u[draw & 1];
}

#endif /* PT_PHILOX_H */
