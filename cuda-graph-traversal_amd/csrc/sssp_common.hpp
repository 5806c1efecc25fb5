// Label arithmetic shared by the single-root and batched delta-stepping
// kernels (reference: src/traversal.cu:211-240 relax, :693-694 bucket bounds).
// Both kernels relax with a 32-bit atomic min on a distance-bits probe and
// commit (distance, parent) labels with a 64-bit atomic min.
#pragma once
#include "common.hpp"

namespace g500 {

constexpr u32 kInfBits = 0x7F800000u;
constexpr u64 kInfLabel = ((u64)kInfBits << 32) | 0xFFFFFFFFull;

// Label = (f32 distance bits << 32) | parent. Non-negative floats order like
// their bit patterns, so comparing the high words compares distances.
__device__ __forceinline__ f32 label_dist(u64 l) { return __uint_as_float((u32)(l >> 32)); }

// The reference's next bucket bound, fl(hi + delta), forced to progress.
__device__ __forceinline__ f32 next_bound(f32 hi, f32 delta) {
    const f32 x = hi + delta;
    return x > hi ? x : __uint_as_float(__float_as_uint(hi) + 1);
}

// split[v] = number of entries of row v with weight < delta (rows are sorted
// by weight), enqueued on `stream` (0: the library stream).
void compute_split(const u64 *rs, const f32 *w, u64 n, f32 delta, u32 *split, hipStream_t stream = nullptr);

}  // namespace g500
