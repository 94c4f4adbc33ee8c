// Host-side number theory for plan construction: 64-bit modular arithmetic
// on unsigned __int128, deterministic Miller-Rabin, and a small u32-limb
// big integer for M = prod p_j and the CRT constants mod q.  Init-time only.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <vector>

namespace nc {
namespace host {

using u128 = unsigned __int128;

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)(((u128)a * b) % p); }

inline uint64_t powmod(uint64_t b, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p;
    b %= p;
    while (e) {
        if (e & 1) r = mulmod(r, b, p);
        b = mulmod(b, b, p);
        e >>= 1;
    }
    return r;
}

inline uint64_t invmod(uint64_t a, uint64_t p) { return powmod(a, p - 2, p); }  // p prime

// floor(w 2^64 / p) for w < p
inline uint64_t shoup(uint64_t w, uint64_t p) { return (uint64_t)(((u128)w << 64) / p); }

// -p^-1 mod 2^64 by Newton iteration (p odd)
inline uint64_t mont_neg_inv(uint64_t p) {
    uint64_t x = 1;
    for (int i = 0; i < 7; ++i) x *= 2 - p * x;
    return 0 - x;
}

// Deterministic for all 64-bit n: the first twelve primes as bases.
inline bool is_prime(uint64_t n) {
    if (n < 2) return false;
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t s : small)
        if (n % s == 0) return n == s;
    uint64_t d = n - 1;
    int r = 0;
    while (!(d & 1)) { d >>= 1; ++r; }
    for (uint64_t a : small) {
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int i = 1; i < r; ++i) {
            x = mulmod(x, x, n);
            if (x == n - 1) { comp = false; break; }
        }
        if (comp) return false;
    }
    return true;
}

// Big unsigned integer, little-endian u32 limbs, fixed capacity.
struct Big {
    std::vector<uint32_t> v;
    explicit Big(size_t limbs = 0) : v(limbs, 0u) {}

    static Big from_u64(uint64_t x, size_t limbs) {
        Big b(limbs);
        b.v[0] = (uint32_t)x;
        if (limbs > 1) b.v[1] = (uint32_t)(x >> 32);
        return b;
    }
    int bits() const {
        for (int i = (int)v.size() - 1; i >= 0; --i)
            if (v[i]) return i * 32 + 32 - __builtin_clz(v[i]);
        return 0;
    }
    bool bit(int i) const { return (i >> 5) < (int)v.size() && ((v[i >> 5] >> (i & 31)) & 1u); }
    void mul_u64(uint64_t f) {  // *this *= f (capacity must suffice)
        uint32_t f0 = (uint32_t)f, f1 = (uint32_t)(f >> 32);
        std::vector<uint32_t> r(v.size(), 0u);
        uint64_t carry = 0;
        for (size_t i = 0; i < v.size(); ++i) {
            uint64_t t = (uint64_t)v[i] * f0 + r[i] + carry;
            r[i] = (uint32_t)t;
            carry = t >> 32;
        }
        if (carry) throw std::runtime_error("Big::mul_u64 overflow");
        carry = 0;
        for (size_t i = 0; i + 1 < v.size(); ++i) {
            uint64_t t = (uint64_t)v[i] * f1 + r[i + 1] + carry;
            r[i + 1] = (uint32_t)t;
            carry = t >> 32;
        }
        if (carry || (f1 && v.back())) throw std::runtime_error("Big::mul_u64 overflow");
        v.swap(r);
    }
    uint64_t mod_u64(uint64_t p) const {
        uint64_t r = 0;
        for (int i = (int)v.size() - 1; i >= 0; --i) r = (uint64_t)((((u128)r) << 32 | v[i]) % p);
        return r;
    }
};

inline int cmp(const Big& a, const Big& b) {
    size_t n = a.v.size() > b.v.size() ? a.v.size() : b.v.size();
    for (size_t i = n; i-- > 0;) {
        uint32_t x = i < a.v.size() ? a.v[i] : 0u, y = i < b.v.size() ? b.v[i] : 0u;
        if (x != y) return x > y ? 1 : -1;
    }
    return 0;
}

// a -= b (requires a >= b); capacity of a
inline void sub_in(Big& a, const Big& b) {
    uint64_t borrow = 0;
    for (size_t i = 0; i < a.v.size(); ++i) {
        uint64_t y = (uint64_t)(i < b.v.size() ? b.v[i] : 0u) + borrow;
        uint64_t x = a.v[i];
        a.v[i] = (uint32_t)(x - y);
        borrow = x < y;
    }
    if (borrow) throw std::runtime_error("Big::sub underflow");
}

// remainder a mod m (bit-serial restoring division)
inline Big mod(const Big& a, const Big& m) {
    size_t n = m.v.size() + 1;
    Big r(n);
    for (int i = a.bits() - 1; i >= 0; --i) {
        uint32_t in = a.bit(i);
        for (size_t l = 0; l < n; ++l) {
            uint32_t nx = (r.v[l] << 1) | in;
            in = r.v[l] >> 31;
            r.v[l] = nx;
        }
        if (cmp(r, m) >= 0) sub_in(r, m);
    }
    r.v.resize(m.v.size());
    return r;
}

// floor(2^e / m)
inline Big pow2_div(int e, const Big& m, size_t limbs) {
    Big q(limbs);
    size_t n = m.v.size() + 1;
    Big r(n);
    for (int i = e; i >= 0; --i) {
        uint32_t in = (i == e);
        for (size_t l = 0; l < n; ++l) {
            uint32_t nx = (r.v[l] << 1) | in;
            in = r.v[l] >> 31;
            r.v[l] = nx;
        }
        if (cmp(r, m) >= 0) {
            sub_in(r, m);
            if ((size_t)(i >> 5) >= limbs) throw std::runtime_error("pow2_div: quotient too wide");
            q.v[i >> 5] |= 1u << (i & 31);
        }
    }
    return q;
}

}  // namespace host
}  // namespace nc
