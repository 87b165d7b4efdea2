// Wavefront reductions for gfx950 (64 lanes) on the cross-lane data paths of
// the VALU, without LDS traffic: DPP moves within rows of 16 lanes, a
// ds_swizzle across the two rows of each 32-lane half, and a readlane of
// each half. Every lane returns the same bits, and the order of operations
// is fixed (it does not depend on the lane or the launch):
//
//   v_i <- v_i op v_(i xor 1)            DPP quad_perm [1,0,3,2]
//   v_i <- v_i op v_(i xor 2)            DPP quad_perm [2,3,0,1]
//   v_i <- v_i op v_(mirror in 8)        DPP row_half_mirror (lane 7-i of its 8)
//   v_i <- v_i op v_(mirror in 16)       DPP row_mirror (lane 15-i of its 16)
//   v_i <- v_i op v_(i xor 16)           ds_swizzle, bit mode, xor mask 16
//   result = v_0 op v_32                 readlane of lanes 0 and 32
//
// After the first two steps the four lanes of a quad hold the same value, so
// the mirror steps pair each lane with a lane of the other quad (of the other
// 8) holding that group's value, and both lanes of every pair compute
// a op b and b op a, which are the same bits for + and for fmax. All 64 lanes
// must be active.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace psim::dev {

constexpr int kDppQuadXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppRowMirror = 0x140;      // lane 15 - i within each row of 16
constexpr int kDppRowHalfMirror = 0x141;  // lane 7 - i within each half-row of 8
// ds_swizzle offset, bit mode: and mask 0x1f, or mask 0, xor mask 0x10
constexpr int kSwizzleXor16 = (0x10 << 10) | 0x1f;

__device__ __forceinline__ uint32_t lo_word(double v) {
    return static_cast<uint32_t>(__builtin_bit_cast(uint64_t, v));
}
__device__ __forceinline__ uint32_t hi_word(double v) {
    return static_cast<uint32_t>(__builtin_bit_cast(uint64_t, v) >> 32);
}
__device__ __forceinline__ double from_words(uint32_t lo, uint32_t hi) {
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// Both words of v moved by the same DPP control (every source lane in range).
template <int kCtrl> __device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(lo_word(v)), kCtrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(hi_word(v)), kCtrl, 0xf, 0xf, false);
    return from_words(static_cast<uint32_t>(lo), static_cast<uint32_t>(hi));
}

__device__ __forceinline__ double swizzle_xor16(double v) {
    const int lo = __builtin_amdgcn_ds_swizzle(static_cast<int>(lo_word(v)), kSwizzleXor16);
    const int hi = __builtin_amdgcn_ds_swizzle(static_cast<int>(hi_word(v)), kSwizzleXor16);
    return from_words(static_cast<uint32_t>(lo), static_cast<uint32_t>(hi));
}

__device__ __forceinline__ double read_lane(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(lo_word(v)), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(hi_word(v)), lane);
    return from_words(static_cast<uint32_t>(lo), static_cast<uint32_t>(hi));
}

template <class Op> __device__ __forceinline__ double wave_reduce(double v, Op op) {
    v = op(v, dpp_move<kDppQuadXor1>(v));
    v = op(v, dpp_move<kDppQuadXor2>(v));
    v = op(v, dpp_move<kDppRowHalfMirror>(v));
    v = op(v, dpp_move<kDppRowMirror>(v));
    v = op(v, swizzle_xor16(v));
    return op(read_lane(v, 0), read_lane(v, 32));
}

__device__ __forceinline__ double wave_sum(double v) {
    return wave_reduce(v, [](double a, double b) { return a + b; });
}

__device__ __forceinline__ double wave_max(double v) {
    return wave_reduce(v, [](double a, double b) { return fmax(a, b); });
}

}  // namespace psim::dev
