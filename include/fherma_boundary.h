/*
 * fherma_boundary.h -- the C++ boundary of the polymul/negacyclic@1.0.0
 * challenge, declared here so that this library's shim exports exactly the
 * symbols the platform's generated harness calls.
 *
 * Types and entry points replace, one for one, the reference's
 * fherma.h:13-47 (Tensor, Point, Inputs, Outputs; fherma_init, fherma_run,
 * fherma_free), named in fherma.toml:4-7 and called by main.cpp:60, 72, 84.
 * The layouts must stay identical to the reference's so that its unchanged
 * main.cpp links against libfherma_negacyclic.so: namespace, member order
 * and member types are part of the C++ ABI.
 *
 * Semantics (solve.cu:1150-1567, kept):
 *  - fherma_init(point): validates N (power of two, 2..65536) and q (odd,
 *    point.q.data.size() == point.L, L <= 32), builds the plan (primes,
 *    tables, self tests) and returns an opaque state.  Throws
 *    std::runtime_error on any failure.
 *  - fherma_run(state, in): in.a / in.b hold N*L u32 limbs each (coefficient
 *    i, limb l at i*L + l); returns c = a*b mod (X^N+1, q) with shape {N, L}.
 *    Throws std::runtime_error("a and b must contain N*L limbs") on a size
 *    mismatch, or on a device error.
 *  - fherma_free(state): releases everything; null safe.
 *  - not re-entrant on one state (like the reference's State).
 */
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace fherma {

template <class T>
struct Tensor {
    std::vector<int64_t> shape;
    std::vector<T> data;

    int64_t count() const {
        int64_t n = 1;
        for (int64_t d : shape) n *= d;
        return n;
    }
};

struct Point {
    uint32_t N = 0;
    uint32_t W = 0;
    uint32_t L = 0;
    Tensor<uint32_t> q;
};

struct Inputs {
    Tensor<uint32_t> a;
    Tensor<uint32_t> b;
};

struct Outputs {
    Tensor<uint32_t> c;
};

}  // namespace fherma

void* fherma_init(const fherma::Point& p);
fherma::Outputs fherma_run(void* state, const fherma::Inputs& in);
void fherma_free(void* state);
