// cc_wave.hpp — wave64 and workgroup reductions for gfx950.
//
// A row of 16 lanes is reduced with four DPP butterflies (quad_perm xor 1, quad_perm xor 2, row_half_mirror,
// row_mirror); the four row results are then read with v_readlane into uniform registers. The workgroup minimum
// costs one barrier: each wave publishes its minimum into one half of a double buffer that moves alternate, so a
// slot written at move i + 2 is only rewritten after every wave has passed the barrier of move i + 1 and finished
// reading move i. Every call must be made by all 64 lanes of the wave.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

namespace cc::hip {

// Bit J of x, sign-extended (0 or -1), as one v_bfe_i32. With a constant offset the compiler turns the sbfe builtin
// into a shift pair, and folding that into what follows costs one to two more instructions per use; the register
// descents take this once per vertex slot and move.
template <int J>
__device__ __forceinline__ int SignedBit(uint32_t x) {
  static_assert(J >= 0 && J < 32, "bit index");
  int r;
  asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(r) : "v"(x), "n"(J));
  return r;
}

// f(std::integral_constant<int, 0>{}), ..., f(std::integral_constant<int, N - 1>{}): a loop whose index is a
// compile-time constant in the body.
template <typename F, int... I>
__device__ __forceinline__ void UnrollImpl(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void Unroll(F&& f) {
  UnrollImpl(f, std::make_integer_sequence<int, N>{});
}

template <int kCtrl>
__device__ __forceinline__ uint32_t DppMove(uint32_t x) {
  // old = x: a lane whose source is invalid keeps its own value, which is neutral for min.
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(x), static_cast<int>(x), kCtrl, 0xF,
                                                           0xF, false));
}

constexpr int kDppXor1 = 0xB1;         // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;         // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror
constexpr int kDppMirror = 0x140;      // row_mirror

template <int kCtrl>
__device__ __forceinline__ uint64_t DppMinStep(uint64_t x) {
  const uint32_t lo = DppMove<kCtrl>(static_cast<uint32_t>(x));
  const uint32_t hi = DppMove<kCtrl>(static_cast<uint32_t>(x >> 32));
  const uint64_t other = (static_cast<uint64_t>(hi) << 32) | lo;
  return other < x ? other : x;
}

template <int kCtrl>
__device__ __forceinline__ uint32_t DppMinStep(uint32_t x) {
  // Every lane's source is valid for these four patterns, so bound_ctrl may be set; with it the compiler folds the
  // move into the minimum (one v_min_u32_dpp per step).
  const uint32_t other =
      static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(x), kCtrl, 0xF, 0xF, true));
  return min(other, x);
}

// Minimum over each row of 16 lanes, present in every lane of the row.
template <typename T>
__device__ __forceinline__ T RowMin(T x) {
  x = DppMinStep<kDppXor1>(x);
  x = DppMinStep<kDppXor2>(x);
  x = DppMinStep<kDppHalfMirror>(x);
  x = DppMinStep<kDppMirror>(x);
  return x;
}

template <int kCtrl>
__device__ __forceinline__ uint32_t DppMaxStep(uint32_t x) {
  const uint32_t other =
      static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(x), kCtrl, 0xF, 0xF, true));
  return max(other, x);
}

// Maximum over each row of 16 lanes, present in every lane of the row.
__device__ __forceinline__ uint32_t RowMax(uint32_t x) {
  x = DppMaxStep<kDppXor1>(x);
  x = DppMaxStep<kDppXor2>(x);
  x = DppMaxStep<kDppHalfMirror>(x);
  x = DppMaxStep<kDppMirror>(x);
  return x;
}

__device__ __forceinline__ uint64_t ReadLane(uint64_t x, int lane) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), lane));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x >> 32), lane));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t ReadLane(uint32_t x, int lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), lane));
}

template <typename T>
__device__ __forceinline__ T Min(T a, T b) {
  return b < a ? b : a;
}

// Minimum over the wave, as a wave-uniform value.
template <typename T>
__device__ __forceinline__ T WaveMin(T x) {
  x = RowMin(x);
  return Min(Min(ReadLane(x, 0), ReadLane(x, 16)), Min(ReadLane(x, 32), ReadLane(x, 48)));
}

// The minimum of a wave's 32-bit keys in every lane, with gfx950's lane swaps instead of four v_readlane: after the
// row minimum, v_permlane32_swap pairs lane i with lane i + 32 and v_permlane16_swap row r with row r ^ 1, so every
// lane ends with the minimum of the wave (one v_readfirstlane moves it to a scalar register).
__device__ __forceinline__ uint32_t WaveMinSwapAll(uint32_t x) {
  x = RowMin(x);
  const auto half = __builtin_amdgcn_permlane32_swap(x, x, false, false);
  x = min(static_cast<uint32_t>(half[0]), static_cast<uint32_t>(half[1]));
  const auto rows = __builtin_amdgcn_permlane16_swap(x, x, false, false);
  return min(static_cast<uint32_t>(rows[0]), static_cast<uint32_t>(rows[1]));
}

// Minimum over a workgroup of kWaves waves (kWaves <= 16) from each wave's own minimum, one barrier. `partial` is
// the half of a double buffer of kWaves entries that this move uses.
template <int kWaves, typename T>
__device__ __forceinline__ T BlockMinOfWaves(T key, T* partial) {
  static_assert(kWaves >= 1 && kWaves <= 16 && (kWaves & (kWaves - 1)) == 0, "waves per block: power of two <= 16");
  if constexpr (kWaves == 1) {
    return key;
  } else {
    const int lane = threadIdx.x & 63;
    if (lane == 0) partial[threadIdx.x >> 6] = key;
    __syncthreads();
    T x = partial[lane & (kWaves - 1)];
    x = RowMin(x);
    return ReadLane(x, 0);
  }
}

template <int kWaves, typename T>
__device__ __forceinline__ T BlockMin(T key, T* partial) {
  return BlockMinOfWaves<kWaves>(WaveMin(key), partial);
}

// A value the same in every lane, moved into scalar registers.
__device__ __forceinline__ int64_t Uniform(int64_t x) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x >> 32)));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Wave-uniform index of the calling wave.
__device__ __forceinline__ int WaveIndex() {
  return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
}

}  // namespace cc::hip
