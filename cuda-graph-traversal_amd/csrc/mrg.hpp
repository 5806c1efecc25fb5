// Splittable MRG arithmetic modulo p = 2^31 - 1, usable on host and device.
//
// Follows the semantics of src/graph_generation.cu:9-94 (the MRG
// x_n = (107374182 x_{n-1} + 104480 x_{n-5}) mod p and its 9-word
// jump-ahead transitions). Residues are computed with 64-bit multiply-adds
// and Mersenne folding; every result is fully reduced to [0, p-1], so the
// numbers are identical to any exact implementation of the same recurrence.
#pragma once
#include <hip/hip_runtime.h>

#include "graph_generation.h"

namespace g500 {

#define G500_HD __host__ __device__ __forceinline__

constexpr u32 kP = 0x7FFFFFFFu;
constexpr u64 kA1 = 107374182u;  // multiplier of x_{n-1}
constexpr u64 kA5 = 104480u;     // multiplier of x_{n-5}

// x mod p for any x < 2^64, fully reduced.
G500_HD u32 fold(u64 x) {
    u64 t = (x & kP) + (x >> 31);                 // < 2^34
    u32 r = (u32)(t & kP) + (u32)(t >> 31);       // < 2^31 + 8
    return r >= kP ? r - kP : r;
}

// One MRG step; returns the new z1 (graph_generation.cu:73-85).
G500_HD u32 mrg_next(Mrg_State &s) {
    u32 nz = fold(kA1 * s.z1 + kA5 * s.z5);
    s.z5 = s.z4;
    s.z4 = s.z3;
    s.z3 = s.z2;
    s.z2 = s.z1;
    s.z1 = nz;
    return nz;
}

// Applies a compressed transition {s,t,u,v,w,a,b,c,d} (the matrix layout of
// graph_generation.cu:56-67):
//   o1 = d z1 + Y (s z2 + a z3 + b z4 + c z5)
//   o2 = c z1 + w z2 + Y (s z3 + a z4 + b z5)
//   o3 = b z1 + v z2 + w z3 + Y (s z4 + a z5)
//   o4 = a z1 + u z2 + v z3 + w z4 + Y s z5
//   o5 = s z1 + t z2 + u z3 + v z4 + w z5        (Y = 104480)
// Four products of residues (< 4 (p-1)^2 < 2^64) are summed before a fold.
G500_HD void mrg_apply(const Mrg_Transition_Matrix &m, Mrg_State &st) {
    const u64 z1 = st.z1, z2 = st.z2, z3 = st.z3, z4 = st.z4, z5 = st.z5;
    u32 y1 = fold(m.s * z2 + m.a * z3 + m.b * z4 + m.c * z5);
    u32 y2 = fold(m.s * z3 + m.a * z4 + m.b * z5);
    u32 y3 = fold(m.s * z4 + m.a * z5);
    u32 y4 = fold(m.s * z5);
    u32 o1 = fold(m.d * z1 + kA5 * y1);
    u32 o2 = fold(m.c * z1 + m.w * z2 + kA5 * y2);
    u32 o3 = fold(m.b * z1 + m.v * z2 + m.w * z3 + kA5 * y3);
    u32 o4 = fold((u64)fold(m.a * z1 + m.u * z2 + m.v * z3 + m.w * z4) + kA5 * y4);
    u32 o5 = fold((u64)fold(m.s * z1 + m.t * z2 + m.u * z3 + m.v * z4) + m.w * z5);
    st.z1 = o1;
    st.z2 = o2;
    st.z3 = o3;
    st.z4 = o4;
    st.z5 = o5;
}

}  // namespace g500
