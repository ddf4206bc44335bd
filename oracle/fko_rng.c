/*
 * fko_rng.c -- host restatement of the counter-based generator that fills
 * synthetic inputs (TEST INFRASTRUCTURE ONLY).
 *
 * The reference fills every array with constants (BGW-Kernels/gppKer.cpp:275-291);
 * seeded random inputs have no counterpart there.  The generator is a
 * SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014; mix constants of Stafford's variant 13)
 * keyed by (seed, stream) and evaluated at a linear index, so a value does
 * not depend on fill order and the device fill kernel reproduces it bit for
 * bit.  Uniform doubles take the top 53 bits; U(lo,hi) = fma(hi-lo, u, lo).
 * Complex element j of an array uses indices 2j (re) and 2j+1 (im).
 */
#include "fko.h"

#include <math.h>

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t fko_hash(uint64_t seed, uint64_t stream, uint64_t index)
{
    uint64_t key = mix64(seed * 0x9E3779B97F4A7C15ULL + mix64(stream + 0x632BE59BD9B4E019ULL));
    return mix64(key + (index + 1) * 0x9E3779B97F4A7C15ULL);
}

double fko_uniform(uint64_t seed, uint64_t stream, uint64_t index, double lo, double hi)
{
    double u = (double)(fko_hash(seed, stream, index) >> 11) * 0x1.0p-53;
    return fma(hi - lo, u, lo);
}

void fko_fill_real(double *out, int64_t n, uint64_t seed, uint64_t stream, double lo, double hi)
{
    for (int64_t i = 0; i < n; ++i) out[i] = fko_uniform(seed, stream, (uint64_t)i, lo, hi);
}

void fko_fill_complex(fko_z *out, int64_t n, uint64_t seed, uint64_t stream,
                      double re_lo, double re_hi, double im_lo, double im_hi)
{
    for (int64_t j = 0; j < n; ++j) {
        out[j].re = fko_uniform(seed, stream, (uint64_t)(2 * j), re_lo, re_hi);
        out[j].im = fko_uniform(seed, stream, (uint64_t)(2 * j + 1), im_lo, im_hi);
    }
}
