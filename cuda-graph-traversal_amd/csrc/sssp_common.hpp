// Label arithmetic shared by the single-root and batched delta-stepping
// kernels (reference: src/traversal.cu:211-240 relax, :693-694 bucket bounds).
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

// Strictly lowers the label at *lab to (cand, parent); true if this call did it.
__device__ __forceinline__ bool relax(u64 *lab, f32 cand, u32 parent) {
    const u32 cb = __float_as_uint(cand);
    const u64 want = ((u64)cb << 32) | parent;
    u64 old = *lab;
    while ((u32)(old >> 32) > cb) {
        const u64 prev = atomicCAS((unsigned long long *)lab, (unsigned long long)old, (unsigned long long)want);
        if (prev == old) return true;
        old = prev;
    }
    return false;
}

// relax() that also tells whether the replaced label was unreached (+inf),
// i.e. whether this call reached the vertex for the first time.
__device__ __forceinline__ bool relax_first(u64 *lab, f32 cand, u32 parent, bool &first) {
    const u32 cb = __float_as_uint(cand);
    const u64 want = ((u64)cb << 32) | parent;
    u64 old = *lab;
    while ((u32)(old >> 32) > cb) {
        const u64 prev = atomicCAS((unsigned long long *)lab, (unsigned long long)old, (unsigned long long)want);
        if (prev == old) {
            first = (u32)(old >> 32) >= kInfBits;
            return true;
        }
        old = prev;
    }
    return false;
}

// split[v] = number of entries of row v with weight < delta (rows are sorted
// by weight), enqueued on `stream` (0: the library stream).
void compute_split(const u64 *rs, const f32 *w, u64 n, f32 delta, u32 *split, hipStream_t stream = nullptr);

}  // namespace g500
