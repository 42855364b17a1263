/*
 * orc_rng.c -- random streams (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * orc_mt19937_uniform: the collision jitter of boxTriCollision.cpp:649-659 and
 *   1079-1090: std::mt19937 seeded with gen.seed(1) feeding
 *   std::uniform_real_distribution<double>(-1, 1) as libstdc++ implements it
 *   (generate_canonical<double,53> draws two 32-bit words w0, w1 and returns
 *   (w0 + w1*2^32) / 2^64, clamped below 1; the distribution returns
 *   u*(hi-lo) + lo).  MT19937 itself is the published Matsumoto-Nishimura
 *   generator.  Pinned against std::mt19937 by tests/golden/jitter_seed1.npy.
 *
 * orc_gaussian: the synthetic-input generator (splitmix64 -> Box-Muller), the
 *   same definition the package's scene generator uses; pinned by a fixture.
 */
#include "oracle.h"
#include <math.h>

typedef struct { uint32_t mt[624]; int idx; } mt_state;

static void mt_seed(mt_state *s, uint32_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->idx = 624;
}

static uint32_t mt_next(mt_state *s)
{
    if (s->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = s->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            s->mt[i] = v;
        }
        s->idx = 0;
    }
    uint32_t y = s->mt[s->idx++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

void orc_mt19937_uniform(uint32_t seed, int count, double lo, double hi, double *out)
{
    mt_state s;
    mt_seed(&s, seed);
    for (int i = 0; i < count; ++i) {
        double sum = 0.0, tmp = 1.0;
        for (int k = 0; k < 2; ++k) {
            sum += (double)mt_next(&s) * tmp;
            tmp *= 4294967296.0;
        }
        double u = sum / tmp;
        if (u >= 1.0) u = nextafter(1.0, 0.0);
        out[i] = u * (hi - lo) + lo;
    }
}

uint64_t orc_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform in (0, 1]: top 53 bits, offset by one ulp step so log() is finite */
static double u01(uint64_t *s) { return ((orc_splitmix64(s) >> 11) + 1) * (1.0 / 9007199254740992.0); }

void orc_gaussian(uint64_t seed, int64_t count, double sigma, double *out)
{
    uint64_t s = seed;
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < count; i += 2) {
        double u1 = u01(&s), u2 = u01(&s);
        double r = sqrt(-2.0 * log(u1));
        out[i] = sigma * r * cos(two_pi * u2);
        if (i + 1 < count) out[i + 1] = sigma * r * sin(two_pi * u2);
    }
}
