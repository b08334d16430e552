// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Written from the
// paper's definition: 10 rounds of two 32x32->64 multiplies (multipliers
// 0xD2511F53, 0xCD9E8D57) with a Weyl key schedule (0x9E3779B9, 0xBB67AE85).
// On gfx950 each multiply is one v_mad_u64_u32 giving both halves.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcs {

struct u32x4 {
    uint32_t x, y, z, w;
};

// Round keys pass through an empty volatile asm so that the schedule is
// recomputed with two scalar adds per round inside the walk loop, instead of
// twenty round keys being held live in SGPRs (which spills them to VGPR lanes).
__device__ __forceinline__ uint32_t sched(uint32_t k)
{
    asm volatile("" : "+s"(k));
    return k;
}

// a ^ b ^ k in one gfx950 v_bitop3_b32 (truth table 0x96); k is wave-uniform.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t k)
{
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "s"(k));
    return r;
}

// Philox4x32-10 of counter c under key (k0, k1).
__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 = sched(k0 + 0x9E3779B9u);
            k1 = sched(k1 + 0xBB67AE85u);
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = u32x4{xor3((uint32_t)(p1 >> 32), c.y, k0), (uint32_t)p1,
                  xor3((uint32_t)(p0 >> 32), c.w, k1), (uint32_t)p0};
    }
    return c;
}

// Per-photon part of round 1 for counters (photon, case, step, stream):
// only the step word changes from one step to the next, so M0 * photon and
// the words it feeds are computed once per photon.
struct PhiloxPre {
    uint32_t ck0;   // case ^ k0
    uint32_t h0;    // hi(M0 * photon) ^ stream ^ k1
    uint32_t l0;    // lo(M0 * photon)
};

__device__ __forceinline__ PhiloxPre philox_pre(uint32_t photon, uint32_t case_id, uint32_t stream,
                                                uint32_t k0, uint32_t k1)
{
    const uint64_t p0 = (uint64_t)0xD2511F53u * photon;
    return PhiloxPre{case_id ^ k0, (uint32_t)(p0 >> 32) ^ stream ^ k1, (uint32_t)p0};
}

// Philox4x32-10 of counter (photon, case, step, stream) given its PhiloxPre.
__device__ __forceinline__ u32x4 philox_step(const PhiloxPre &pre, uint32_t step, uint32_t k0, uint32_t k1)
{
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * step;
    u32x4 c{(uint32_t)(p1 >> 32) ^ pre.ck0, (uint32_t)p1, pre.h0, pre.l0};
#pragma unroll
    for (int r = 1; r < 10; ++r) {
        k0 = sched(k0 + 0x9E3779B9u);
        k1 = sched(k1 + 0xBB67AE85u);
        const uint64_t q0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t q1 = (uint64_t)0xCD9E8D57u * c.z;
        c = u32x4{xor3((uint32_t)(q1 >> 32), c.y, k0), (uint32_t)q1,
                  xor3((uint32_t)(q0 >> 32), c.w, k1), (uint32_t)q0};
    }
    return c;
}

// 32 random bits -> float in (0, 1]: (x + 1/2) * 2^-32, rounded to FP32.
__device__ __forceinline__ float u01f(uint32_t x)
{
    return __builtin_fmaf((float)x, 2.3283064365386963e-10f, 1.1641532182693481e-10f);
}

}  // namespace mcs
