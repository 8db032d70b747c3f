// The split of an fp64 (or fp32) complex value into bf16 words for the split-BF16x3
// contraction: hi = bf16(x), lo = bf16(x - hi) per real part, packed as one
// 32-bit word of (re, im) bf16 pairs per part.  Shared by the contraction
// kernels (split at staging) and the recursion steps that write pre-split
// stacks, so both produce the same words.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tbk {

__device__ __forceinline__ uint32_t bf16_bits(float f) {
    return static_cast<uint32_t>(__builtin_bit_cast(unsigned short, static_cast<__bf16>(f)));
}

// (hi, lo) words of one complex value
__device__ __forceinline__ uint2 split_word(double2 v) {
    const float hr = static_cast<float>(v.x), hi_ = static_cast<float>(v.y);
    const __bf16 br = static_cast<__bf16>(hr), bi = static_cast<__bf16>(hi_);
    const float lr = static_cast<float>(v.x - static_cast<double>(static_cast<float>(br)));
    const float li = static_cast<float>(v.y - static_cast<double>(static_cast<float>(bi)));
    uint2 w;
    w.x = static_cast<uint32_t>(__builtin_bit_cast(unsigned short, br)) |
          (static_cast<uint32_t>(__builtin_bit_cast(unsigned short, bi)) << 16);
    w.y = bf16_bits(lr) | (bf16_bits(li) << 16);
    return w;
}

// (hi, lo) words of an fp32 complex value (the fp32 recursion's stacks):
// x - hi is exact in fp32, since hi is x rounded to 8 significant bits
__device__ __forceinline__ uint2 split_word(float2 v) {
    const __bf16 br = static_cast<__bf16>(v.x), bi = static_cast<__bf16>(v.y);
    const float lr = v.x - static_cast<float>(br);
    const float li = v.y - static_cast<float>(bi);
    uint2 w;
    w.x = static_cast<uint32_t>(__builtin_bit_cast(unsigned short, br)) |
          (static_cast<uint32_t>(__builtin_bit_cast(unsigned short, bi)) << 16);
    w.y = bf16_bits(lr) | (bf16_bits(li) << 16);
    return w;
}

// Layout of a stack of split words (N entries of K elements): k-blocked,
// the kWordBlock-element pieces of all entries at one k side by side,
//   word (n, k) at (k / kWordBlock) * bs + n * kWordBlock + k % kWordBlock,
// bs = N * kWordBlock (the block stride).  The contraction reads 16 columns
// of hundreds of entries at a time; blocked, those lie within a few pages
// instead of one page per entry.  An entry's own words start at
// n * kWordBlock; bs = 0 means a plain contiguous entry.
constexpr int kWordBlock = 256;
__host__ __device__ __forceinline__ int64_t word_offset(int64_t k, int64_t bs) {
    return bs ? (k / kWordBlock) * bs + (k % kWordBlock) : k;
}

}  // namespace tbk
