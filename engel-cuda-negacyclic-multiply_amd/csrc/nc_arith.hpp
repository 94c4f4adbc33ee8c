// Modular word arithmetic for the gfx950 kernels, as two policies that the
// NTT code is written against:
//   W62: primes p < 2^62, 64-bit residues, 64x64 products built by the
//        compiler from v_mad_u64_u32 / v_mul_hi_u32 / v_mul_lo_u32
//   W30: primes p < 2^30, 32-bit residues, one v_mul_hi_u32 / v_mul_lo_u32
//        per partial product
// Both keep Harvey's lazy ranges: forward (Cooley-Tukey) values in [0, 4p),
// inverse (Gentleman-Sande) values in [0, 2p), which needs 4p < 2^bits.
// Twiddle products use Shoup's precomputed quotient ws = floor(w 2^bits / p);
// the pointwise product is a Montgomery product (R = 2^bits).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nc_common.hpp"

namespace nc {

using u128 = unsigned __int128;

struct W62 {
    using T = uint64_t;
    using Prime = PrimeConst;
    using Twid = Tw;
    static constexpr int kBits = 64;

    // a*w mod p in [0, 2p) for any a and w < p.
    __device__ __forceinline__ static T shoup(T a, T w, T ws, T p) { return a * w - __umul64hi(a, ws) * p; }
    __device__ __forceinline__ static T sub_if_ge(T x, T m) { return x >= m ? x - m : x; }
    // a b 2^-64 mod p in [0, 2p) for a, b < 2p
    __device__ __forceinline__ static T mont(T a, T b, T p, T mneg) {
        const u128 t = (u128)a * b;
        const uint64_t lo = (uint64_t)t, hi = (uint64_t)(t >> 64);
        return hi + __umul64hi(lo * mneg, p) + (lo != 0);
    }
};

struct W30 {
    using T = uint32_t;
    using Prime = PrimeConst32;
    using Twid = Tw32;
    static constexpr int kBits = 32;

    __device__ __forceinline__ static T shoup(T a, T w, T ws, T p) { return a * w - __umulhi(a, ws) * p; }
    // x - m if x >= m else x: when x < m, x - m wraps above x (m < 2^31), so
    // the minimum picks the right one in two instructions.
    __device__ __forceinline__ static T sub_if_ge(T x, T m) { return min(x, x - m); }
    // a b 2^-32 mod p in (0, 2p) for a, b < 2p: with m = lo p^-1 mod 2^32
    // (p^-1 = -mneg), m p has the low word of a b, so
    // (a b - m p) / 2^32 = hi - hi(m p), exactly, in (-p, p) as 4p^2 < p 2^32.
    // The + p rides in the product's multiply-add (a b + p 2^32 < 2^64).
    __device__ __forceinline__ static T mont(T a, T b, T p, T mneg) {
        uint64_t p_hi = (uint64_t)p << 32;
        asm("" : "+v"(p_hi));  // a register-pair addend of the multiply-add
        const uint64_t t = (uint64_t)a * b + p_hi;
        const uint32_t lo = (uint32_t)t, hi_p = (uint32_t)(t >> 32);
        T pinv = 0u - mneg;
        asm("" : "+v"(pinv));  // keep the inverse: no negate per product
        return hi_p - __umulhi(lo * pinv, p);
    }
};

// Harvey butterflies on a policy A.
template <class A>
struct Bfly {
    using T = typename A::T;
    // CT: x, y in [0, 4p) -> (x + wy, x - wy) in [0, 4p)
    __device__ __forceinline__ static void ct(T& x, T& y, typename A::Twid t, T p, T p2) {
        const T u = A::sub_if_ge(x, p2);
        const T v = A::shoup(y, t.w, t.ws, p);
        x = u + v;
        y = u - v + p2;
    }
    // GS: x, y in [0, 2p) -> (x + y, w (x - y)) in [0, 2p)
    __device__ __forceinline__ static void gs(T& x, T& y, typename A::Twid t, T p, T p2) {
        const T u = x + y;
        const T d = x - y + p2;
        x = A::sub_if_ge(u, p2);
        y = A::shoup(d, t.w, t.ws, p);
    }
    // GS stage 0 with N^-1 R merged: x' = (x + y) nr, y' = (x - y) psi^-N/2 nr
    __device__ __forceinline__ static void gs_last(T& x, T& y, const typename A::Prime& c) {
        const T u = x + y;
        const T d = x - y + c.p2;
        x = A::shoup(u, c.nr, c.nr_s, c.p);
        y = A::shoup(d, c.wr, c.wr_s, c.p);
    }
};

// 30-bit butterflies.  Forward twiddles are stored negated (w- = 2^32 - w,
// quotient ws of w): with q = hi(y ws), nv = q p + y w- = -(y w - q p) = -v
// mod 2^32, v in [0, 2p) (Shoup), so x' = u - nv = u + v and
// y' = u + nv + 2p = u - v + 2p (one three-operand add), both exact in
// [0, 4p): 8 instructions instead of 9.
template <>
struct Bfly<W30> {
    using T = uint32_t;
    __device__ __forceinline__ static void ct(T& x, T& y, Tw32 t, T p, T p2) {
        const T u = W30::sub_if_ge(x, p2);
        const T nv = __umulhi(y, t.ws) * p + y * t.w;
        x = u - nv;
        y = u + nv + p2;
    }
    // a w - hi(a ws) p written as a w + hi(a ws) (2^32 - p): the same value
    // mod 2^32, and the multiply-add form is one v_mad_u64_u32 (3 instructions
    // instead of 4).  The empty asm keeps the compiler from folding 2^32 - p
    // back into a subtraction.
    __device__ __forceinline__ static T shoup(T a, T w, T ws, T p) {
        T np = 0u - p;
        asm("" : "+v"(np));
        return a * w + __umulhi(a, ws) * np;
    }
    __device__ __forceinline__ static void gs(T& x, T& y, Tw32 t, T p, T p2) {
        const T u = x + y;
        const T d = x - y + p2;
        x = W30::sub_if_ge(u, p2);
        y = shoup(d, t.w, t.ws, p);
    }
    __device__ __forceinline__ static void gs_last(T& x, T& y, const PrimeConst32& c) {
        const T u = x + y;
        const T d = x - y + c.p2;
        x = shoup(u, c.nr, c.nr_s, c.p);
        y = shoup(d, c.wr, c.wr_s, c.p);
    }
};

}  // namespace nc
