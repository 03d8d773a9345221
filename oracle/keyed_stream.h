/*
 * TEST INFRASTRUCTURE ONLY (oracle). Nothing in the product may include this file.
 *
 * The keyed sample stream: a counter-based replacement for the reference's
 * random sources (McRng, Banshee/rng/mcrng.h:45-76, and glibc rand() used by
 * GiTracer, Banshee/tracer/gitracer.cpp:75,119). Every random number of a
 * camera sample is a pure function of (seed, camera-sample index, stream,
 * counter), so the compiled reference, the CPU restatement and the GPU path
 * draw identical numbers for identical paths.
 *
 * Mixer: "lowbias32" from Chris Wellons, "Prospecting for Hash Functions"
 * (2018): x ^= x>>16; x *= 0x7feb352d; x ^= x>>15; x *= 0x846ca68b; x ^= x>>16.
 */
#ifndef ORACLE_KEYED_STREAM_H
#define ORACLE_KEYED_STREAM_H

#include <stdint.h>

enum { KS_IMG = 0, KS_LIGHT = 1, KS_BRDF = 2, KS_RAND = 3, KS_PROBE = 4 };

static inline uint32_t ks_mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

/* Per-camera-sample base key. */
static inline uint32_t ks_base(uint32_t seed, uint32_t camsample)
{
    return ks_mix32(ks_mix32(seed + 0x9e3779b9U) ^ camsample);
}

/* counter must stay below 2^24. */
static inline uint32_t ks_u32(uint32_t base, uint32_t stream, uint32_t counter)
{
    return ks_mix32(base ^ ((stream << 24) | counter));
}

/* Same uint->float conversion as McRng::NextFloat (mcrng.h:54-57). */
static inline float ks_float(uint32_t u)
{
    return (float)u * 2.328306e-10f;
}

/* glibc rand() stand-in: 31-bit value in [0, RAND_MAX]. */
static inline int ks_rand(uint32_t base, uint32_t counter)
{
    return (int)(ks_u32(base, KS_RAND, counter) >> 1);
}

#endif
