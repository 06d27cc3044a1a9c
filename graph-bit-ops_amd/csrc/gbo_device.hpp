// Device-side helpers for gfx950 (CDNA4, 64-lane wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "gbo_internal.hpp"

namespace gbo {

constexpr int kWave = 64;

template <int BS> struct WordOf;
template <> struct WordOf<8> { using type = uint8_t; };
template <> struct WordOf<16> { using type = uint16_t; };
template <> struct WordOf<32> { using type = uint32_t; };
template <> struct WordOf<64> { using type = uint64_t; };
template <int BS> using word_t = typename WordOf<BS>::type;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Index of the calling wavefront in the grid.  threadIdx.x / 64 is the same
// in every lane, but the compiler cannot prove it; readfirstlane makes the
// value (and what is derived from it) scalar, so per-chunk loop bounds live
// in SGPRs and their branches are uniform.
__device__ __forceinline__ int wave_index() {
  return __builtin_amdgcn_readfirstlane((int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave));
}

__device__ __forceinline__ int popcw(uint64_t x) { return __popcll(x); }
__device__ __forceinline__ int popcw(uint32_t x) { return __popc(x); }
__device__ __forceinline__ int popcw(uint16_t x) { return __popc((uint32_t)x); }
__device__ __forceinline__ int popcw(uint8_t x) { return __popc((uint32_t)x); }

// Reverses the low BS bits of x (bit 0 <-> bit BS-1).
template <int BS>
__device__ __forceinline__ uint64_t brev_low(uint64_t x) {
  return __brevll(x) >> (64 - BS);
}

// OR of v over the wave (every lane gets it, as a wave-uniform value): DPP
// rotations by 1, 2, 4, 8 inside each 16-lane row, then the four rows'
// results read from one lane each.  Every lane of the wave must be active.
__device__ __forceinline__ uint32_t wave_or32(uint32_t v) {
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x121, 0xF, 0xF, false);  // row_ror:1
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x122, 0xF, 0xF, false);  // row_ror:2
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false);  // row_ror:4
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);  // row_ror:8
  return (uint32_t)(__builtin_amdgcn_readlane((int)v, 0) | __builtin_amdgcn_readlane((int)v, 16) |
                    __builtin_amdgcn_readlane((int)v, 32) | __builtin_amdgcn_readlane((int)v, 48));
}
__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
  return (uint64_t)wave_or32((uint32_t)v) | ((uint64_t)wave_or32((uint32_t)(v >> 32)) << 32);
}

// Inclusive prefix sum of v over the 64 lanes with DPP moves (no LDS round
// trips): shifts by 1, 2, 4, 8 inside each 16-lane row, then lane 15 of
// rows 0 and 2 is added to rows 1 and 3 and lane 31 to rows 2 and 3.  A
// lane whose DPP source falls outside its row (or whose row is masked off)
// reads the `old` operand, 0.  Every lane of the wave must be active.
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15 -> rows 1, 3
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31 -> rows 2, 3
  return v;
}

// *flag = 1 if any lane of the wave passes `any`.  Termination flags are
// only ever set to the same value, so this is a plain store after a plain
// read instead of a read-modify-write: tens of thousands of waves do not
// queue on one address.
__device__ __forceinline__ void wave_flag(bool any, unsigned long long* flag) {
  if (__ballot(any) && (threadIdx.x & (kWave - 1)) == 0 && *flag == 0ull) *flag = 1ull;
}

// Read of p[i] for a wave-uniform i from memory no thread writes during the
// kernel: through a constant-address-space pointer the load is a scalar
// (s_load) one, off the vector memory path.
template <typename T>
__device__ __forceinline__ T ld_uniform(const T* p, int64_t i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// OR of the 64/BS groups of BS bits of a ballot: bit r is set when any lane
// with lane % BS == r had its bit set (a row's hit over the lane groups that
// hold it), by scalar shifts instead of a cross-lane reduction.
template <int BS>
__device__ __forceinline__ uint64_t fold_groups(uint64_t b) {
  if constexpr (BS <= 32) b |= b >> 32;
  if constexpr (BS <= 16) b |= b >> 16;
  if constexpr (BS <= 8) b |= b >> 8;
  if constexpr (BS == 64) return b;
  else return b & ((1ull << BS) - 1ull);
}

// Mask with the low BS bits set.
template <int BS>
__device__ __forceinline__ uint64_t low_mask() {
  if constexpr (BS == 64) return ~0ull;
  else return (1ull << BS) - 1;
}

// Atomic OR of a BS-bit word into a word array (bytes and shorts go through
// the aligned 32-bit word that contains them; little-endian).  Returns the
// previous value of the BS-bit word.
template <int BS>
__device__ __forceinline__ uint64_t atomic_or_word(word_t<BS>* base, int64_t idx, uint64_t v) {
  if constexpr (BS == 64) {
    return atomicOr((unsigned long long*)(base + idx), (unsigned long long)v);
  } else if constexpr (BS == 32) {
    return atomicOr((unsigned int*)(base + idx), (unsigned int)v);
  } else {
    const uintptr_t addr = (uintptr_t)(base + idx);
    unsigned int* w32 = (unsigned int*)(addr & ~(uintptr_t)3);
    const int shift = (int)(addr & 3) * 8;
    const unsigned int old = atomicOr(w32, (unsigned int)(v << shift));
    return (old >> shift) & (unsigned int)low_mask<BS>();
  }
}

// Reductions over the lanes of one group of BS lanes, or across the G
// groups that hold the same row index (xor offsets BS, 2BS, ...).
// 32-bit values reduce inside a 16-lane row with DPP moves (no LDS round
// trip): quad_perm swaps of 1 and 2, then row_half_mirror (lane i <-> 7-i of
// each 8) and row_mirror (i <-> 15-i), which pair every lane with the other
// half once its quad / half already holds the minimum.  Every lane of the
// wave must be active.
template <int BS, class T>
__device__ __forceinline__ T group_min(T v) {
  if constexpr (sizeof(T) == 4) {
    auto step = [&](T o) { v = o < v ? o : v; };
    step((T)__builtin_amdgcn_update_dpp((int)0, (int)v, 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
    if constexpr (BS > 2) step((T)__builtin_amdgcn_update_dpp((int)0, (int)v, 0x4E, 0xF, 0xF, false));  // [2,3,0,1]
    if constexpr (BS > 4) step((T)__builtin_amdgcn_update_dpp((int)0, (int)v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    if constexpr (BS > 8) step((T)__builtin_amdgcn_update_dpp((int)0, (int)v, 0x140, 0xF, 0xF, false));  // row_mirror
    for (int off = 16; off < BS; off <<= 1) {
      T o = __shfl_xor(v, off);
      v = o < v ? o : v;
    }
  } else {
    for (int off = 1; off < BS; off <<= 1) {
      T o = __shfl_xor(v, off);
      v = o < v ? o : v;
    }
  }
  return v;
}

}  // namespace gbo

#define GBO_HIP_CHECK(expr)                                                              \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return gbo::set_error(GBO_EDEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #expr,      \
                            hipGetErrorString(_e));                                      \
  } while (0)

#define GBO_LAUNCH_CHECK() GBO_HIP_CHECK(hipGetLastError())
