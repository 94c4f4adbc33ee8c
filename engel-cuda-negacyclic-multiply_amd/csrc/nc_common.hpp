// Structures shared by the host plan builder and the gfx950 kernels.
#pragma once

#include <cstdint>

namespace nc {

constexpr int kMaxLimbs = 32;     // L <= 32 (solve.cu:1170)
constexpr int kMaxPrimes = 48;    // k <= 48 (solve.cu:22)
constexpr int kMaxLogN = 16;      // N <= 65536
constexpr int kMaxLogBlock = 14;  // largest NTT block held in one workgroup's LDS
constexpr int kMaxOuterBits = 4;  // outer NTT passes cover at most 4 stages

// Bytes of the per-device scratch array of nc_multiply_device_noalloc
// (residue rows of a and b): one product at N = 65536, four at N = 16384.
constexpr unsigned long long kScratchBytes = 32ull << 20;

// Per-prime constants for the NTT / pointwise kernels.
struct PrimeConst {
    uint64_t p;         // prime, p < 2^62, p = 1 mod 2N
    uint64_t p2;        // 2p
    uint64_t mont_neg;  // -p^-1 mod 2^64 (Montgomery REDC constant)
    uint64_t nr;        // N^-1 * 2^64 mod p: undoes 1/N and the pointwise 2^-64
    uint64_t nr_s;      // Shoup companion floor(nr 2^64 / p)
    uint64_t wr;        // psi^-1 * N^-1 * 2^64 mod p: last GS stage twiddle, merged
    uint64_t wr_s;
    uint64_t pad;
};

// Per-prime constants of the CRT (fast CRT, solve.cu:752-836).
struct CrtConst {
    uint64_t y;      // (M / p_j)^-1 mod p_j
    uint64_t y_s;    // Shoup companion
    uint64_t mag_hi; // floor(2^128 / p_j) = mag_hi 2^64 + mag_lo
    uint64_t mag_lo;
};

// Modulus q and its reduction constants, limbs zero padded to kMaxLimbs.
struct QConst {
    uint32_t q[kMaxLimbs];
    uint32_t neg_m[kMaxLimbs];  // (-M) mod q, M = prod p_j
    uint32_t mu[4];             // floor(2^(bits(q) - 1 + 72) / q), 3 limbs used
    uint32_t q_bits;
    uint32_t L;
    uint32_t pad[2];
};

// Twiddle pair: value and Shoup companion floor(w 2^64 / p).
struct Tw {
    uint64_t w;
    uint64_t ws;
};

// ---- 30-bit prime set (p < 2^30, 32-bit words) ----
constexpr int kMaxPrimes30 = 72;

// The 30-bit constants are kMergeVariants30 blocks of k rows: block v has nr
// and wr times 2^v, for inner blocks that merge v levels of the transform
// into the pointwise product (their output is 2^-v scaled).
struct PrimeConst32 {
    uint32_t p, p2;
    uint32_t mont_neg;  // -p^-1 mod 2^32
    uint32_t nr, nr_s;  // N^-1 2^32 (2^v) mod p, Shoup companion floor(nr 2^32 / p)
    uint32_t wr, wr_s;  // psi^-bitrev(1) N^-1 2^32 (2^v) mod p
    uint32_t c32;       // 2^32 mod p (decomposition fold constant)
};
constexpr int kMergeVariants30 = 3;

struct Tw32 {
    uint32_t w;
    uint32_t ws;  // floor(w 2^32 / p)
};

// Byte-split int8 MFMA decomposition table (30-bit primes, L % 4 == 0):
// row 4 j + u of tile j / 8 holds the balanced base-256 digit u (int8) of
// (2^(8 m) 2^32 mod p_j) for m < 4L, and 0 for m >= 4L; rows of primes
// >= k are 0.  A per-row int32 bias undoes the kernel's -128 byte offset
// (see decompose30_mfma_kernel).
constexpr int kMfmaK = 128;                        // contraction length (bytes of a coefficient, padded)
constexpr int kMfmaRowStride = kMfmaK + 16;        // LDS row pitch: conflict-free 16-byte reads
constexpr int kMfmaTiles = (kMaxPrimes30 + 7) / 8; // 8 primes x 4 byte chunks = 32 rows per tile

// Byte-split int8 MFMA CRT table (30-bit primes).  Contraction index
// kappa = 4 j + u (prime j, byte u of v_j; chunks of 8 primes = 32 bytes).
// Output row (lambda, b) sits at 32 (lambda / 8) + 8 ((lambda % 8) / 2) +
// 4 (lambda % 2) + b, so lane half h of v_mfma_i32_32x32x32_i8 receives whole
// pseudo-limbs lambda = 8 rt + 2 m + h (m = 0..3) of row tile rt.  With RT row
// tiles and LQ = min(8 RT - 4, 32): lambda < LQ are the limbs of
// sum_j v_j ((M / p_j) mod q), lambda = 8 RT - 4, 8 RT - 3 the two limbs of
// the fraction sum_j v_j floor(2^56 / p_j), lambda = 8 RT - 2, 8 RT - 1 (byte
// 0) a row of ones that yields the byte sum of the coefficient's v_j.
constexpr int kCrtMaxKC = (kMaxPrimes30 + 7) / 8;   // contraction chunks (8 primes each)
constexpr int kCrtRowStride = kCrtMaxKC * 32 + 16;  // LDS row pitch: conflict-free 16-byte reads
constexpr int kCrtMaxRT = 5;                        // (32 + 4) pseudo-limbs / 8 per row tile

// Montgomery finish of the MFMA CRT (R = 2^32): the table carries
// (M / p_j) mod q times R, so the limb sum x satisfies x = c R (mod q).
struct CrtMfmaConst {
    uint32_t neg_m_r[kMaxLimbs];  // ((-M) mod q) R mod q
    uint32_t q_inv;               // -q^-1 mod 2^32
    uint32_t pad[3];
};

// CRT constants for a 30-bit prime: y = (M/p)^-1 mod p (Shoup pair) and
// g = floor(2^96 / p) = g_hi 2^64 + g_lo for the fixed-point fraction
// floor(v 2^64 / p) = floor(v g / 2^32) (at most one below).
struct CrtConst32 {
    uint32_t y, y_s;
    uint32_t g_hi, pad;
    uint64_t g_lo;
};

}  // namespace nc
