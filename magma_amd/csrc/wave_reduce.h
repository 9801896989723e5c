// wave_reduce.h -- wavefront-wide integer max / min and the (key, position)
// arg-max of the pivot searches.
//
// A reduction is a prefix scan within each row of 16 lanes by DPP (row_shr 1,
// 2, 4, 8: one integer max or min per step, the shifted-in lanes reading the
// identity) followed by the four row ends (lanes 15, 31, 47, 63) read with
// readlane and combined on the scalar unit.  Pivot keys are the bit patterns
// of non-negative floating values plus one (0 marks a row that is not a
// candidate): unsigned comparison of those orders them as the values, so the
// arg-max needs integer steps only, where a (float or double key, position)
// pair needed a compare, a tie test and two selects per step.  The arg-max is
// the largest key, ties to the smallest position: idamax's first maximum when
// positions are the rows' current logical positions.
#pragma once

#include <climits>

namespace magma_amd {

__device__ __forceinline__ unsigned wave_max_u32(unsigned v)
{
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false));   // row_shr:1
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false));   // row_shr:2
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false));   // row_shr:4
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false));   // row_shr:8
    const unsigned a = __builtin_amdgcn_readlane(v, 15), b = __builtin_amdgcn_readlane(v, 31);
    const unsigned c = __builtin_amdgcn_readlane(v, 47), d = __builtin_amdgcn_readlane(v, 63);
    return max(max(a, b), max(c, d));
}

__device__ __forceinline__ int wave_min_i32(int v)
{
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x111, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x112, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x114, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x118, 0xF, 0xF, false));
    const int a = __builtin_amdgcn_readlane(v, 15), b = __builtin_amdgcn_readlane(v, 31);
    const int c = __builtin_amdgcn_readlane(v, 47), d = __builtin_amdgcn_readlane(v, 63);
    return min(min(a, b), min(c, d));
}

// ordered pivot key of |a| (a non-negative, +inf allowed)
__device__ __forceinline__ unsigned long long pivot_key(double a) { return (unsigned long long)__double_as_longlong(a) + 1ull; }
__device__ __forceinline__ unsigned pivot_key(float a) { return (unsigned)__float_as_int(a) + 1u; }

// (key, pos) over the wavefront: largest key, ties to the smallest pos.  The
// result is wave-uniform.  When one lane alone holds the largest high word
// (the usual case: equal high words need equal exponents and 20 leading
// mantissa bits) that lane is the arg-max and the low-word and position
// passes are skipped.
__device__ __forceinline__ void wave_argmax_key(unsigned long long &key, int &pos)
{
    const unsigned hi = (unsigned)(key >> 32), lo = (unsigned)key;
    const unsigned mh = wave_max_u32(hi);
    const unsigned long long tie = __builtin_amdgcn_ballot_w64(hi == mh);
    if (__builtin_popcountll(tie) == 1) {
        const int l = (int)__builtin_ctzll(tie);
        pos = __builtin_amdgcn_readlane(pos, l);
        key = ((unsigned long long)mh << 32) | (unsigned)__builtin_amdgcn_readlane((int)lo, l);
        return;
    }
    const unsigned ml = wave_max_u32(hi == mh ? lo : 0u);
    pos = wave_min_i32((hi == mh && lo == ml) ? pos : INT_MAX);
    key = ((unsigned long long)mh << 32) | ml;
}

__device__ __forceinline__ void wave_argmax_key(unsigned &key, int &pos)
{
    const unsigned m = wave_max_u32(key);
    const unsigned long long tie = __builtin_amdgcn_ballot_w64(key == m);
    if (__builtin_popcountll(tie) == 1) {
        pos = __builtin_amdgcn_readlane(pos, (int)__builtin_ctzll(tie));
    } else {
        pos = wave_min_i32(key == m ? pos : INT_MAX);
    }
    key = m;
}

// The same within groups of G lanes (8, 16, 32 or 64; several matrices per
// wavefront): quad swaps, the half-row and row mirrors, then the row values
// by readlane.  Every lane gets its group's result.
template <int G>
__device__ __forceinline__ unsigned group_max_u32(unsigned v, int lane)
{
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));    // quad_perm [1,0,3,2]
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));    // quad_perm [2,3,0,1]
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));   // row_half_mirror
    if constexpr (G >= 16) v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));
    if constexpr (G >= 32) {
        const unsigned r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
        const unsigned r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
        if constexpr (G == 64) v = max(max(r0, r1), max(r2, r3));
        else v = lane < 32 ? max(r0, r1) : max(r2, r3);
    }
    return v;
}

template <int G>
__device__ __forceinline__ int group_min_i32(int v, int lane)
{
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0xB1, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x4E, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x141, 0xF, 0xF, false));
    if constexpr (G >= 16) v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x140, 0xF, 0xF, false));
    if constexpr (G >= 32) {
        const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
        const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
        if constexpr (G == 64) v = min(min(r0, r1), min(r2, r3));
        else v = lane < 32 ? min(r0, r1) : min(r2, r3);
    }
    return v;
}

template <int G>
__device__ __forceinline__ void group_argmax_key(unsigned long long &key, int &pos, int lane)
{
    const unsigned hi = (unsigned)(key >> 32), lo = (unsigned)key;
    const unsigned mh = group_max_u32<G>(hi, lane);
    if constexpr (G >= 32) {
        // one lane per group with the largest high word: that lane is the
        // arg-max (as in wave_argmax_key)
        const unsigned long long tie = __builtin_amdgcn_ballot_w64(hi == mh);
        const unsigned t0 = (unsigned)tie, t1 = (unsigned)(tie >> 32);
        if (G == 64 ? __builtin_popcountll(tie) == 1 : (__builtin_popcount(t0) == 1 && __builtin_popcount(t1) == 1)) {
            const int l0 = (G == 64) ? (int)__builtin_ctzll(tie) : (int)__builtin_ctz(t0);
            const int l1 = (G == 64) ? l0 : 32 + (int)__builtin_ctz(t1);
            const int p0 = __builtin_amdgcn_readlane(pos, l0), p1 = __builtin_amdgcn_readlane(pos, l1);
            const unsigned q0 = __builtin_amdgcn_readlane((int)lo, l0), q1 = __builtin_amdgcn_readlane((int)lo, l1);
            pos = lane < 32 ? p0 : p1;
            key = ((unsigned long long)mh << 32) | (lane < 32 ? q0 : q1);
            return;
        }
    }
    const unsigned ml = group_max_u32<G>(hi == mh ? lo : 0u, lane);
    pos = group_min_i32<G>((hi == mh && lo == ml) ? pos : INT_MAX, lane);
    key = ((unsigned long long)mh << 32) | ml;
}

}  // namespace magma_amd
