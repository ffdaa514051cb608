// Counter-based random numbers. The reference draws from the process-global
// glibc rand() seeded with time(NULL) (lib/nmath/prng.inl:54-83), so no run of
// it is reproducible; here every draw is a pure function of
// (seed, job, pixel, sample, query index, draw index), which makes a render
// independent of how its rays are batched, ordered or interleaved with other
// jobs. Mixer: the "lowbias32" integer hash (Chris Wellons, public domain).
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define XT_HD __host__ __device__ __forceinline__
#else
#define XT_HD inline
#endif

XT_HD uint32_t xt_mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Key of one camera sample of one job.
XT_HD uint32_t xt_path_key(uint32_t seed, uint32_t job, uint32_t pixel, uint32_t sample) {
    uint32_t h = xt_mix32(seed + 0x9e3779b9u);
    h = xt_mix32(h ^ job);
    h = xt_mix32(h ^ pixel);
    return xt_mix32(h ^ (sample * 0x85ebca6bu + 0x27d4eb2fu));
}

// Stream of the draws made while processing query `q` of a path (q = 0 is
// the camera ray itself, i.e. lens and anti-aliasing draws).
XT_HD uint32_t xt_stream(uint32_t key, uint32_t q) { return xt_mix32(key ^ (q * 0xc2b2ae35u + 0x165667b1u)); }

XT_HD uint32_t xt_draw(uint32_t stream, uint32_t i) { return xt_mix32(stream + i * 0x9e3779b9u + 0x7f4a7c15u); }

// Uniform in [0, 1], both ends included like prng_c(0, 1).
XT_HD double xt_unit64(uint32_t x) { return (double)x * (1.0 / 4294967295.0); }
XT_HD float xt_unit32(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777215.0f); }
