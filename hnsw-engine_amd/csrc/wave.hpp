// Wavefront-level helpers for CDNA4 (wave64). One wavefront owns one query in the search
// kernels, so "wave-uniform" values live in scalar registers and LDS traffic is ordered with
// wave-scope fences instead of workgroup barriers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HM_DEV __device__ __forceinline__

namespace hm {

constexpr int kWave = 64;

HM_DEV int lane_id() { return __lane_id(); }

// Orders this wave's LDS/global accesses before and after the call (the rocPRIM pattern).
HM_DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The active lanes whose p holds, as a mask: the condition's own lane mask (HIP's __ballot first
// copies the condition into a vector register and compares it again).
HM_DEV uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Masks of the active lanes whose v compares below a wave-uniform s, straight from the compare
// instruction.
HM_DEV uint64_t lanes_lt_u32(uint32_t v, uint32_t s) { return __builtin_amdgcn_uicmp(v, s, 36); }
HM_DEV uint64_t lanes_le_u32(uint32_t v, uint32_t s) { return __builtin_amdgcn_uicmp(v, s, 37); }
HM_DEV uint64_t lanes_lt_u64(uint64_t v, uint64_t s) { return __builtin_amdgcn_uicmpl(v, s, 36); }

HM_DEV int popc(uint64_t m) { return __popcll(m); }

// Number of set bits of `m` below this lane.
HM_DEV int rank_below(uint64_t m) {
    uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
    return __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
}

HM_DEV int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

HM_DEV uint32_t uniform_u32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

HM_DEV uint64_t uniform_u64(uint64_t v) {
    uint32_t lo = uniform_u32((uint32_t)v), hi = uniform_u32((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

HM_DEV int shfl(int v, int src) { return __shfl(v, src, kWave); }
HM_DEV uint32_t shfl_u32(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, kWave); }
HM_DEV uint64_t shfl_u64(uint64_t v, int src) {
    uint32_t lo = shfl_u32((uint32_t)v, src), hi = shfl_u32((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}
HM_DEV int shfl_xor(int v, int m) { return __shfl_xor(v, m, kWave); }

// Sum over aligned groups of g lanes (g a wave-uniform power of two, 1..64); every lane of a
// group receives its group's sum. Inside a row of 16 lanes the exchanges are DPP operands
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror: after each step all
// lanes of the current group hold equal values, so a mirror reads the other half's sum), taken
// or not by a select instead of a branch; the 32- and 64-lane steps cross rows through
// ds_bpermute. All lanes must be active.
HM_DEV int group_sum(int v, int g) {
    int t = __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true);
    v += g >= 2 ? t : 0;
    t = __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true);
    v += g >= 4 ? t : 0;
    t = __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, true);
    v += g >= 8 ? t : 0;
    t = __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, true);
    v += g >= 16 ? t : 0;
    if (g >= 32) {
        v += shfl_xor(v, 16);
        if (g >= 64) v += shfl_xor(v, 32);
    }
    return v;
}

// Float sum over each row of 16 lanes, every lane of the row receiving it (as bits): the same
// DPP exchanges as group_sum with g = 16.
HM_DEV int group_sum_f32(float x) {
    int v = __float_as_int(x);
    v = __float_as_int(__int_as_float(v) +
                       __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true)));
    v = __float_as_int(__int_as_float(v) +
                       __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true)));
    v = __float_as_int(__int_as_float(v) +
                       __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, true)));
    v = __float_as_int(__int_as_float(v) +
                       __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, true)));
    return v;
}

// Minimum over the wave (every lane receives it): DPP within rows of 16, then the four rows.
HM_DEV uint32_t wave_min_u32(uint32_t v) {
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, true));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, true));
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
    const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    return min(min(r0, r1), min(r2, r3));
}

HM_DEV int wave_sum(int v) {
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor(v, m);
    return v;
}

HM_DEV int wave_max(int v) {
    for (int m = 32; m >= 1; m >>= 1) v = max(v, shfl_xor(v, m));
    return v;
}

// Minimum (key, id) pair over the wave; every lane receives it.
HM_DEV void wave_min_pair(uint64_t &k, uint32_t &id) {
    for (int m = 32; m >= 1; m >>= 1) {
        uint32_t klo = (uint32_t)__shfl_xor((int)(uint32_t)k, m, kWave);
        uint32_t khi = (uint32_t)__shfl_xor((int)(uint32_t)(k >> 32), m, kWave);
        uint32_t oid = (uint32_t)__shfl_xor((int)id, m, kWave);
        uint64_t ok = ((uint64_t)khi << 32) | klo;
        if (ok < k || (ok == k && oid < id)) {
            k = ok;
            id = oid;
        }
    }
}

}  // namespace hm
