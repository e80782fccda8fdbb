/*
 * trk_philox.h -- counter-based random numbers for per-history streams.
 *
 * The reference draws every random number from the compiler runtime's single
 * shared `random_number` stream (43 call sites in Monte_Carlo.f90, 8 in
 * Cross_sections.f90; seeding at MPI_subroutines.f90:387-400), so its
 * histories cannot be reproduced or replayed in another order.  Here every
 * particle owns a Philox4x32-10 stream (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11), keyed by the run
 * seed and addressed by (iteration, particle id, draw index).  A GPU that
 * processes particles in any order and a CPU that processes them in the
 * reference's global time order therefore consume identical numbers.
 *
 * Draw n of a particle uses block n>>1 of its stream:
 *   ctr = {n>>1, id_lo, id_hi, iteration}, key = {seed_lo, seed_hi}
 * and takes words (0,1) for even n, (2,3) for odd n, combined into a 53-bit
 * integer u; the value is (u + 0.5) * 2^-53, which lies strictly inside (0,1)
 * so -log(RN) is always finite.
 */
#ifndef TRK_PHILOX_H
#define TRK_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define TRK_RFN __host__ __device__ static inline
#else
#define TRK_RFN static inline
#endif

typedef struct { uint32_t v[4]; } trk_u32x4;

TRK_RFN uint32_t trk_mulhilo32(uint32_t a, uint32_t b, uint32_t *hi)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}

/* Philox4x32 with 10 rounds: multipliers and Weyl key increments as given in
 * the SC'11 paper (Table 2 and section 5). */
TRK_RFN trk_u32x4 trk_philox4x32_10(trk_u32x4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, hi1;
        uint32_t lo0 = trk_mulhilo32(M0, c.v[0], &hi0);
        uint32_t lo1 = trk_mulhilo32(M1, c.v[2], &hi1);
        trk_u32x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

/* Uniform double in (0,1) for draw `n` of particle `id` in iteration `iter`. */
TRK_RFN double trk_uniform(uint64_t seed, uint64_t id, uint32_t iter, uint32_t n)
{
    trk_u32x4 c;
    c.v[0] = n >> 1;
    c.v[1] = (uint32_t)id;
    c.v[2] = (uint32_t)(id >> 32);
    c.v[3] = iter;
    trk_u32x4 r = trk_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t a = (n & 1u) ? r.v[2] : r.v[0];
    uint32_t b = (n & 1u) ? r.v[3] : r.v[1];
    uint64_t u = ((((uint64_t)a) << 32) | (uint64_t)b) >> 11;
    return ((double)u + 0.5) * 0x1.0p-53;
}

/* Identity of a child particle: a bijective 64-bit mix (splitmix64 finaliser,
 * Steele/Lea/Flood, OOPSLA'14) of the parent id, the parent's event number and
 * the slot of the child within that event. */
TRK_RFN uint64_t trk_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

TRK_RFN uint64_t trk_child_id(uint64_t parent, uint32_t event_no, uint32_t slot)
{
    return trk_mix64(parent * 0x9E3779B97F4A7C15ull + (((uint64_t)event_no << 3) | (uint64_t)slot) + 1ull);
}

#endif /* TRK_PHILOX_H */
