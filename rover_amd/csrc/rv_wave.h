// rv_wave.h -- 64-lane wavefront exchange and sorting helpers for gfx950.
//
// Lane exchanges inside a quad use DPP quad permutes (a VALU modifier, no
// LDS round trip); exchanges across 4..16 lanes use ds_swizzle in bitmask
// mode (lane i reads lane i ^ j inside its 32-lane group, no address
// register); the 32-lane exchange uses v_permlane32_swap.
#ifndef RV_WAVE_H
#define RV_WAVE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

// v_permlane32_swap with v in both operands exchanges the upper 32 lanes of
// the first with the lower 32 lanes of the second: the first result holds
// lanes 0..31 in both halves, the second lanes 32..63 in both halves.
// Value of lane (lane & 31): the lower half copied to the upper one.
__device__ __forceinline__ uint32_t rv_low_half(uint32_t v) {
    return __builtin_amdgcn_permlane32_swap(v, v, false, false)[0];
}

// Value of lane (lane ^ 32).
__device__ __forceinline__ uint32_t rv_swap_halves(uint32_t v) {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return __builtin_amdgcn_inverse_ballot_w64(0xffffffffull) ? r[1] : r[0];
}

// Value of lane - 1 inside each 16-lane row; the first lane of a row
// receives 0 (DPP row_shr:1, a VALU operand modifier).
__device__ __forceinline__ float rv_row_prev(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xF, 0xF, false));
}

// Value of lane (lane ^ J) for J in {1, 2, 4, 8, 16, 32}.
template <int J>
__device__ __forceinline__ uint32_t rv_xor_lane(uint32_t v) {
    static_assert(J == 1 || J == 2 || J == 4 || J == 8 || J == 16 || J == 32, "xor");
    if (J == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    if (J == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    if (J == 32) return rv_swap_halves(v);
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (J << 10));
}

template <int J>
__device__ __forceinline__ uint64_t rv_xor_lane64(uint64_t v) {
    const uint32_t lo = rv_xor_lane<J>((uint32_t)v);
    const uint32_t hi = rv_xor_lane<J>((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Lanes that keep the smaller key in the bitonic step (K, J): those where
// (lane & K) == 0 and (lane & J) == 0 agree.
template <int K, int J>
constexpr uint64_t rv_keep_min_mask() {
    uint64_t m = 0;
    for (int l = 0; l < 64; l++)
        if (((l & K) == 0) == ((l & J) == 0)) m |= 1ull << l;
    return m;
}

// One compare-exchange step of a bitonic sort: partner lane ^ J, blocks of
// K lanes ascending where (lane & K) == 0.  The keep-smaller/keep-larger
// choice is a compile-time lane mask combined with the compare's lane mask
// on the scalar unit; one select per 32-bit half applies it.  Must be
// executed by all 64 lanes.
template <int K, int J>
__device__ __forceinline__ uint64_t rv_bitonic_step(uint64_t key) {
    const uint64_t other = rv_xor_lane64<J>(key);
    constexpr uint64_t keep_min = rv_keep_min_mask<K, J>();
    const uint64_t lt = __ballot(key < other);
    const bool take_key = __builtin_amdgcn_inverse_ballot_w64(~(lt ^ keep_min));
    return take_key ? key : other;
}

// Bitonic stages k = 2..16: blocks of 16 lanes sorted, alternately
// ascending (lane & 16 == 0) and descending.
__device__ __forceinline__ uint64_t rv_sort_blocks16(uint64_t key) {
    key = rv_bitonic_step<2, 1>(key);
    key = rv_bitonic_step<4, 2>(key);
    key = rv_bitonic_step<4, 1>(key);
    key = rv_bitonic_step<8, 4>(key);
    key = rv_bitonic_step<8, 2>(key);
    key = rv_bitonic_step<8, 1>(key);
    key = rv_bitonic_step<16, 8>(key);
    key = rv_bitonic_step<16, 4>(key);
    key = rv_bitonic_step<16, 2>(key);
    key = rv_bitonic_step<16, 1>(key);
    return key;
}

// Ascending sort of one 64-bit key per lane inside each 32-lane half
// (lanes 0..31 and 32..63 are sorted independently; K = 64 makes the last
// merge ascending in both halves).
__device__ __forceinline__ uint64_t rv_sort_halves(uint64_t key) {
    key = rv_sort_blocks16(key);
    key = rv_bitonic_step<64, 16>(key);
    key = rv_bitonic_step<64, 8>(key);
    key = rv_bitonic_step<64, 4>(key);
    key = rv_bitonic_step<64, 2>(key);
    key = rv_bitonic_step<64, 1>(key);
    return key;
}

// Ascending sort of one 64-bit key per lane across the whole wavefront.
__device__ __forceinline__ uint64_t rv_sort_wave(uint64_t key) {
    key = rv_sort_blocks16(key);
    key = rv_bitonic_step<32, 16>(key);
    key = rv_bitonic_step<32, 8>(key);
    key = rv_bitonic_step<32, 4>(key);
    key = rv_bitonic_step<32, 2>(key);
    key = rv_bitonic_step<32, 1>(key);
    key = rv_bitonic_step<64, 32>(key);
    key = rv_bitonic_step<64, 16>(key);
    key = rv_bitonic_step<64, 8>(key);
    key = rv_bitonic_step<64, 4>(key);
    key = rv_bitonic_step<64, 2>(key);
    key = rv_bitonic_step<64, 1>(key);
    return key;
}

// Value of lane - 1 (lane 0 receives its own value), via DPP wave_shr:1.
__device__ __forceinline__ uint64_t rv_prev_lane64(uint64_t v) {
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)v, (int)(uint32_t)v, 0x138,
                                               0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(v >> 32),
                                               (int)(uint32_t)(v >> 32), 0x138, 0xF, 0xF,
                                               false);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

#endif
