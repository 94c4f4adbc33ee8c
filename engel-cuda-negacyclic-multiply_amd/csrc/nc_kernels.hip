// gfx950 kernels of the negacyclic product c = a*b mod (X^N + 1, q).
//
// Pipeline per product (reference: solve.cu:1489-1528):
//   decompose   limbs -> residues mod k primes        (solve.cu:533-570)
//   NTT product forward negacyclic NTT of a and b, pointwise product,
//               inverse NTT                           (solve.cu:1006-1044, 620-642)
//   CRT         residues -> L limbs mod q             (solve.cu:752-836)
// for one of two prime sets (nc_arith.hpp): 62-bit primes in 64-bit words
// (the reference's prime rule) or 30-bit primes in 32-bit words (about twice
// as many primes, each butterfly a handful of 32-bit instructions).
//
// NTT formulation: negacyclic NTT with the psi twist merged into the
// butterflies (Cooley-Tukey forward, natural -> bit-reversed order;
// Gentleman-Sande inverse, bit-reversed -> natural), so there is no bit
// reversal pass and no separate twist / untwist pass.  Stage s (0-based)
// pairs indices that differ in bit beta = logN - 1 - s; its twiddle for the
// group i = x >> (beta + 1) is tw[2^s + i], tw[x] = psi^bitrev(x) (forward)
// or psi^-bitrev(x) (inverse).  N^-1 and the Montgomery factor of the
// pointwise product are folded into the last inverse stage.
//
// Work decomposition of the NTT: a prime row of N values is split into
// 2^s1 contiguous blocks of 2^b values (b <= 14, so a block fits one
// workgroup's LDS).  The first s1 forward stages (bits >= b) and the last s1
// inverse stages run in "outer" kernels where each thread owns one column of
// 2^s1 values; all remaining stages run in the "inner" kernel, one block per
// (sub-)workgroup, 16 values per thread in registers, LDS for the exchanges
// between radix-16 passes.  The inner kernel also does the pointwise product,
// so forward-inner, pointwise and inverse-inner share one HBM round trip.
// The outer stages can ride in neighbours instead of their own pass: the
// matrix-core decomposition runs s1 <= 2 forward stages, and the
// matrix-core CRT runs the last two inverse stages at s1 = 2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "nc_arith.hpp"
#include "nc_common.hpp"
#include "nc_kernels.hpp"

namespace nc {

// Workgroups go round-robin over the 8 XCDs (id % 8).  Give each XCD one
// contiguous range of logical ids so that neighbouring work items (same
// prime, hence same twiddle rows) share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t g, uint32_t G) {
    const uint32_t per = G >> 3;
    if (g >= (per << 3)) return g;
    return (g & 7u) * per + (g >> 3);
}

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// ===========================================================================
// decomposition: limbs -> canonical residues
// A workgroup stages 64 coefficients' limbs in LDS once (coalesced) and its
// four waves cover all k primes (wave w takes primes w, w+4, ...), so each
// limb is read from HBM once for all primes.
// ===========================================================================
template <int LT>
__device__ __forceinline__ void stage_limbs(uint32_t* tile, const uint32_t* src, uint32_t cnt, uint32_t L) {
    constexpr int STRIDE = LT + 1;  // odd stride: conflict-free LDS reads per coefficient
    const uint32_t words = cnt * L;
    for (uint32_t w = threadIdx.x; w < words; w += 256) {
        const uint32_t c = w / L, l = w - c * L;
        tile[c * STRIDE + l] = src[w];
    }
}

// 62-bit primes: r = sum_l limb_l (2^(32 l) 2^64 mod p) < 2^99 exactly in 128
// bits, residue = REDC(r) = sum_l limb_l 2^(32 l) mod p.
template <int LT>
__global__ __launch_bounds__(256) void decompose62_kernel(const uint32_t* __restrict__ a, uint64_t* __restrict__ res_a,
                                                          uint32_t na, const uint32_t* __restrict__ b,
                                                          uint64_t* __restrict__ res_b,
                                                          const uint64_t* __restrict__ place,
                                                          const PrimeConst* __restrict__ pc, uint32_t N, uint32_t L,
                                                          uint32_t k) {
    constexpr int TILE = 64, STRIDE = LT + 1;
    __shared__ uint32_t tile[TILE * STRIDE];
    const uint32_t tiles = (N + TILE - 1) / TILE;
    const uint32_t poly = blockIdx.x / tiles;
    const uint32_t base = (blockIdx.x % tiles) * TILE;
    const uint32_t cnt = N - base < (uint32_t)TILE ? N - base : (uint32_t)TILE;
    const bool is_a = poly < na;
    const uint32_t pi = is_a ? poly : poly - na;
    stage_limbs<LT>(tile, (is_a ? a : b) + ((size_t)pi * N + base) * L, cnt, L);
    uint64_t* dst = (is_a ? res_a : res_b) + (size_t)pi * k * N + base;
    __syncthreads();
    const uint32_t c = threadIdx.x & 63u;
    const uint32_t wave = uniform(threadIdx.x >> 6);
    if (c >= cnt) return;
    uint32_t limb[LT];
#pragma unroll
    for (int l = 0; l < LT; ++l) limb[l] = (uint32_t)l < L ? tile[c * STRIDE + l] : 0u;
    for (uint32_t j = wave; j < k; j += 4) {
        const uint64_t* pl = place + (size_t)j * kMaxLimbs;
        u128 s = 0;
#pragma unroll
        for (int l = 0; l < LT; ++l) s += (u128)limb[l] * pl[l];
        const uint64_t p = pc[j].p;
        const uint64_t lo = (uint64_t)s;
        const uint64_t r = (uint64_t)(s >> 64) + __umul64hi(lo * pc[j].mont_neg, p) + (lo != 0);
        dst[(size_t)j * N + c] = W62::sub_if_ge(r, p);
    }
}

// 30-bit primes: S accumulates limb_l (2^(32 l) 2^32 mod p) (each < 2^62)
// two at a time in 64 bits, folded back below 2^62 + 2^32 with
// S = hi(S) (2^32 mod p) + lo(S) after every pair; a final REDC (R = 2^32)
// gives sum_l limb_l 2^(32 l) mod p.
template <int LT>
__global__ __launch_bounds__(256) void decompose30_kernel(const uint32_t* __restrict__ a, uint32_t* __restrict__ res_a,
                                                          uint32_t na, const uint32_t* __restrict__ b,
                                                          uint32_t* __restrict__ res_b,
                                                          const uint32_t* __restrict__ place,
                                                          const PrimeConst32* __restrict__ pc, uint32_t N, uint32_t L,
                                                          uint32_t k) {
    constexpr int TILE = 64, STRIDE = LT + 1;
    __shared__ uint32_t tile[TILE * STRIDE];
    const uint32_t tiles = (N + TILE - 1) / TILE;
    const uint32_t poly = blockIdx.x / tiles;
    const uint32_t base = (blockIdx.x % tiles) * TILE;
    const uint32_t cnt = N - base < (uint32_t)TILE ? N - base : (uint32_t)TILE;
    const bool is_a = poly < na;
    const uint32_t pi = is_a ? poly : poly - na;
    stage_limbs<LT>(tile, (is_a ? a : b) + ((size_t)pi * N + base) * L, cnt, L);
    uint32_t* dst = (is_a ? res_a : res_b) + (size_t)pi * k * N + base;
    __syncthreads();
    const uint32_t c = threadIdx.x & 63u;
    const uint32_t wave = uniform(threadIdx.x >> 6);
    if (c >= cnt) return;
    uint32_t limb[LT];
#pragma unroll
    for (int l = 0; l < LT; ++l) limb[l] = (uint32_t)l < L ? tile[c * STRIDE + l] : 0u;
    for (uint32_t j = wave; j < k; j += 4) {
        const uint32_t* pl = place + (size_t)j * kMaxLimbs;
        const PrimeConst32 cj = pc[j];
        uint64_t s = 0;
#pragma unroll
        for (int l = 0; l < LT; l += 2) {
            s += (uint64_t)limb[l] * pl[l];
            if (l + 1 < LT) s += (uint64_t)limb[l + 1] * pl[l + 1];
            s = (uint64_t)(uint32_t)(s >> 32) * cj.c32 + (uint32_t)s;
        }
        s = (uint64_t)(uint32_t)(s >> 32) * cj.c32 + (uint32_t)s;  // < 2^61 + 2^32 < p 2^32
        const uint32_t lo = (uint32_t)s;
        const uint32_t r = (uint32_t)(s >> 32) + __umulhi(lo * cj.mont_neg, cj.p) + (lo != 0);
        dst[(size_t)j * N + c] = W30::sub_if_ge(r, cj.p);
    }
}

// 30-bit primes on the matrix cores: the digit sum is an int8 contraction.
// With a_m the 4L bytes of a coefficient and t_um the balanced base-256
// digits (in [-128, 127]) of T_m = 2^(8m) 2^32 mod p_j (table rows 4j + u),
//   V_j = sum_m a_m T_m = sum_u 2^(8u) X_u,   X_u = sum_m a_m t_um.
// The MFMA computes X_u = sum_m (a_m - 128) t_um + 128 sum_m t_um (signed
// int8 inputs, the row bias 128 sum_m t_um as accumulator input, exact int32
// sums), so V_j = sum_u 2^(8u) X_u = lo + 2^16 hi with lo = X_0 + 2^8 X_1,
// hi = X_2 + 2^8 X_3 (|X_u| < 2^22: both fit int32).  0 <= V_j < 4L 255 2^30
// < p 2^32, and one REDC (R = 2^32) gives the residue (the table carries the
// 2^32).
// Layout (v_mfma_i32_32x32x32_i8): the A operand is the table (rows =
// prime x byte, 8 primes per 32-row tile), the B operand is 32 coefficients'
// bytes, so lane l receives all four byte sums of primes 2h + l/32 (h = 0..3)
// for coefficient l % 32 and recombines them without any lane exchange.
typedef int nc_v4i __attribute__((ext_vector_type(4)));
typedef int nc_v16i __attribute__((ext_vector_type(16)));

constexpr int kMfmaCoeffTiles = 4;  // 32-coefficient tiles per wave

template <class A, int R, int NE>
__device__ __forceinline__ void ct_pass(typename A::T (&v)[NE], uint32_t gx0, int lo, int log_n,
                                        const typename A::Twid* __restrict__ tw, typename A::T p, typename A::T p2);

// S1 > 0 also runs the first S1 forward NTT stages (the outer stages that
// otherwise take their own pass over the residue rows): the wave's four tiles
// are then 4 / 2^S1 column groups of 2^S1 tiles spaced N / 2^S1 apart, so each
// lane holds, per prime, the 2^S1 values those stages combine.  Output is
// then lazy in [0, 4p), exactly what ntt_outer_fwd_kernel<W30, S1> leaves.
// WS > 1 (launches with few work items): WS waves share each chunk, wave w
// of a group taking prime tiles [w T / WS, (w + 1) T / WS) of the T tiles,
// so each wave's serial chain is WS times shorter; a work item then holds
// 4 / WS chunks.
// Copies n 16-byte pieces from global memory to LDS (a 256-thread
// workgroup, at most MAXPER pieces per thread): every load is issued before
// the first store, and no load sits in a branch (a thread past the end
// re-reads piece n - 1), so the copy costs about one memory round trip.
template <int MAXPER, class Src, class Dst>
__device__ __forceinline__ void copy16_to_lds(uint32_t n, Src src_of, Dst dst_of) {
    nc_v4i t[MAXPER];
#pragma unroll
    for (int i = 0; i < MAXPER; ++i) {
        const uint32_t w = threadIdx.x + 256u * i;
        t[i] = *src_of(w < n ? w : n - 1);
    }
#pragma unroll
    for (int i = 0; i < MAXPER; ++i) {
        const uint32_t w = threadIdx.x + 256u * i;
        if (w < n) *dst_of(w) = t[i];
    }
}

struct DecompPrimeLds {
    uint32_t p, mont_neg;
};

// LDS of the matrix-core decomposition: the int8 table, each prime's
// constants, the row biases and, for the G = 2^S1 spread tiles of the fused
// stages, psi^bitrev(x), x = 1 .. G - 1, of each prime.
template <int G>
struct DecompLds {
    int8_t tab[kMfmaTiles * 32 * kMfmaRowStride];
    DecompPrimeLds prime[kMfmaTiles * 8];
    __attribute__((aligned(16))) int32_t bias[kMfmaTiles * 32];
    Tw32 tw[kMfmaTiles * 8][G];
};

// Loads the decomposition's tables into s (every thread of a 256-thread
// workgroup takes part; the caller synchronises before the first item).
template <int S1>
__device__ __forceinline__ void decompose30_mfma_tables(DecompLds<1 << S1>& s, const int8_t* __restrict__ tab,
                                                        const int32_t* __restrict__ bias,
                                                        const PrimeConst32* __restrict__ pc,
                                                        const Tw32* __restrict__ twf, uint32_t k, uint32_t tiles,
                                                        uint32_t tw_log) {
    constexpr int G = 1 << S1;
    constexpr int kPieces = kMfmaTiles * 32 * (kMfmaK / 16);
    copy16_to_lds<(kPieces + 255) / 256>(
        tiles * 32 * (kMfmaK / 16),
        [&](uint32_t w) { return reinterpret_cast<const nc_v4i*>(tab + (size_t)w * 16); },
        [&](uint32_t w) {
            return reinterpret_cast<nc_v4i*>(s.tab + (w / (kMfmaK / 16)) * kMfmaRowStride + (w % (kMfmaK / 16)) * 16);
        });
    for (uint32_t r = threadIdx.x; r < tiles * 32; r += 256) s.bias[r] = bias[r];
    for (uint32_t j = threadIdx.x; j < tiles * 8; j += 256) {
        s.prime[j] = j < k ? DecompPrimeLds{pc[j].p, pc[j].mont_neg} : DecompPrimeLds{1u, 0u};
        if constexpr (S1 > 0) {
#pragma unroll
            for (int x = 1; x < G; ++x) s.tw[j][x] = j < k ? twf[((size_t)j << tw_log) + x] : Tw32{0u, 0u};
        }
    }
}

// One work item of the decomposition: 4 NG / WS chunks of 32 columns of one
// polynomial (src: its N x L limbs; dst: its [k][N] residue rows), starting
// at chunk chunk0.  No barrier inside.
template <int S1, int WS>
__device__ __forceinline__ void decompose30_mfma_item(const DecompLds<1 << S1>& s, const uint32_t* __restrict__ src,
                                                      uint32_t* __restrict__ dst, uint32_t chunk0, uint32_t N,
                                                      uint32_t L, uint32_t k, uint32_t tiles) {
    constexpr int G = 1 << S1;                 // spread tiles per column group
    constexpr int NG = kMfmaCoeffTiles / G;    // column groups per wave
    static_assert(NG >= 1, "at most 4 spread tiles");
    static_assert(WS == 1 || WS == 2 || WS == 4, "waves per chunk");
    const int8_t* lds_tab = s.tab;
    const DecompPrimeLds* lds_prime = s.prime;
    const int32_t* lds_bias = s.bias;
    const uint32_t span = N >> S1;                      // columns of one spread tile group
    const uint32_t chunks = span / 32;                  // 32-column chunks per polynomial
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t col = lane & 31u, half = lane >> 5;
    const uint32_t nbytes = 4 * L;
#pragma unroll 1
    for (int g = 0; g < NG; ++g) {
        const uint32_t ch = chunk0 + (wave / WS) * NG + g;
        if (ch >= chunks) break;  // wave-uniform; no barrier in the item loop
        const uint32_t c0 = ch * 32 + col;  // this lane's column; tile e holds coefficient c0 + e span
        nc_v4i bf[G][kMfmaK / 32];
#pragma unroll
        for (int e = 0; e < G; ++e) {
            const char* cp = reinterpret_cast<const char*>(src + (size_t)(c0 + e * span) * L);
#pragma unroll
            for (int ks = 0; ks < kMfmaK / 32; ++ks) {
                // bytes m >= 4L meet zero table entries, so any bytes serve
                // there: read the coefficient's last 16 instead.  The load
                // stays unconditional, and all of them issue before one wait
                // (a guarded load waits for itself inside its branch).
                const uint32_t off = 32 * ks + 16 * half;
                const uint32_t offc = off < nbytes ? off : nbytes - 16;
                bf[e][ks] = *reinterpret_cast<const nc_v4i*>(cp + offc) ^ (int)0x80808080;  // bytes a_m - 128
            }
        }
        uint32_t t_lo = 0, t_hi = tiles;
        if constexpr (WS > 1) {
            t_lo = (wave % WS) * tiles / WS;
            t_hi = (wave % WS + 1) * tiles / WS;
        }
        for (uint32_t t = t_lo; t < t_hi; ++t) {
            uint32_t v[4][G];  // [h][e]
            // 2^16 in a register the compiler cannot see through, so that
            // lo + 2^16 hi is one v_mad_i64_i32 rather than 64-bit shifts and adds
            int32_t two16;
            asm("s_mov_b32 %0, 0x10000" : "=s"(two16));
            // accumulator input: output row 8 h + 4 half + u of the tile is
            // table row (prime 2 h + half, byte u)
            nc_v16i c_in;
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const nc_v4i b4 = *reinterpret_cast<const nc_v4i*>(lds_bias + t * 32 + 8 * h + 4 * half);
#pragma unroll
                for (int u = 0; u < 4; ++u) c_in[4 * h + u] = b4[u];
            }
#pragma unroll
            for (int e = 0; e < G; ++e) {
                const int8_t* arow = lds_tab + (t * 32 + col) * kMfmaRowStride + 16 * half;
                nc_v16i acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(*reinterpret_cast<const nc_v4i*>(arow), bf[e][0],
                                                                    c_in, 0, 0, 0);
#pragma unroll
                for (int ks = 1; ks < kMfmaK / 32; ++ks)
                    acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(*reinterpret_cast<const nc_v4i*>(arow + 32 * ks),
                                                                bf[e][ks], acc, 0, 0, 0);
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    const DecompPrimeLds cj = lds_prime[t * 8 + 2 * h + half];
                    const int32_t lo = acc[4 * h] + (int32_t)((uint32_t)acc[4 * h + 1] << 8);
                    const int32_t hi = acc[4 * h + 2] + (int32_t)((uint32_t)acc[4 * h + 3] << 8);
                    const uint64_t V = (uint64_t)((int64_t)hi * two16 + (int64_t)lo);
                    // REDC: V + m p = 0 mod 2^32, (V + m p) / 2^32 < 2p
                    const uint32_t m = (uint32_t)V * cj.mont_neg;
                    const uint32_t r = (uint32_t)(((uint64_t)m * cj.p + V) >> 32);
                    // canonical for plain residues; the fused CT stages take [0, 4p) as is
                    v[h][e] = S1 > 0 ? r : W30::sub_if_ge(r, cj.p);
                }
            }
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const uint32_t j = t * 8 + 2 * h + half;
                if constexpr (S1 > 0) {
                    const uint32_t p = lds_prime[j].p;
                    ct_pass<W30, S1, G>(v[h], c0, (int)(31 - __builtin_clz(span)), (int)(31 - __builtin_clz(N)),
                                        &s.tw[j][0], p, 2 * p);
                }
                if (j < k) {
#pragma unroll
                    for (int e = 0; e < G; ++e) dst[(size_t)j * N + c0 + e * span] = v[h][e];
                }
            }
        }
    }
}

template <int S1, int WS = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void decompose30_mfma_kernel(
    const uint32_t* __restrict__ a, uint32_t* __restrict__ res_a, uint32_t na, const uint32_t* __restrict__ b,
    uint32_t* __restrict__ res_b, const int8_t* __restrict__ tab, const int32_t* __restrict__ bias,
    const PrimeConst32* __restrict__ pc, const Tw32* __restrict__ twf, uint32_t N, uint32_t L, uint32_t k,
    uint32_t tiles, uint32_t nb, uint32_t tw_log) {
    constexpr int NG = kMfmaCoeffTiles / (1 << S1);
    __shared__ DecompLds<1 << S1> s;
    decompose30_mfma_tables<S1>(s, tab, bias, pc, twf, k, tiles, tw_log);
    __syncthreads();
    const uint32_t chunks = (N >> S1) / 32;  // 32-column chunks per polynomial
    constexpr uint32_t kChunksPerItem = 4 * NG / WS;
    const uint32_t wgs_per_poly = (chunks + kChunksPerItem - 1) / kChunksPerItem;
    // persistent: the tables above are loaded once per workgroup and serve
    // every work item it takes (work item = 4 NG / WS chunks of one polynomial)
    for (uint32_t item = blockIdx.x; item < wgs_per_poly * (na + nb); item += gridDim.x) {
        const uint32_t poly = item / wgs_per_poly;
        const uint32_t chunk0 = (item - poly * wgs_per_poly) * kChunksPerItem;
        const bool is_a = poly < na;
        const uint32_t pi = is_a ? poly : poly - na;
        const uint32_t* src = (is_a ? a : b) + (size_t)pi * N * L;
        uint32_t* dst = (is_a ? res_a : res_b) + (size_t)pi * k * N;
        decompose30_mfma_item<S1, WS>(s, src, dst, chunk0, N, L, k, tiles);
    }
}

// ===========================================================================
// NTT passes on registers.  A pass handles R consecutive index bits
// [lo, lo + R) of the NE values a thread holds; register e sits at global
// row index gx0 + (e << lo).  When NE > 2^R the bits of e above R are spare:
// they index further groups the pass transforms side by side (only a pass
// with lo = 0 has them, and the low bits of gx0 are then zero), so stage kk
// of register e uses twiddle (e >> (kk + 1)) of the stage's run.
// ===========================================================================
template <class A, int R, int NE>
__device__ __forceinline__ void ct_pass(typename A::T (&v)[NE], uint32_t gx0, int lo, int log_n,
                                        const typename A::Twid* __restrict__ tw, typename A::T p,
                                        typename A::T p2) {
#pragma unroll
    for (int kk = R - 1; kk >= 0; --kk) {
        const int beta = lo + kk;
        const uint32_t base = (1u << (log_n - 1 - beta)) + (gx0 >> (beta + 1));
        constexpr int MAXH = NE / 2;
        typename A::Twid t[MAXH];
        const int nh = NE >> (kk + 1);
#pragma unroll
        for (int h = 0; h < MAXH; ++h)
            if (h < nh) t[h] = tw[base + h];
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            if (e & (1 << kk)) continue;
            Bfly<A>::ct(v[e], v[e | (1 << kk)], t[e >> (kk + 1)], p, p2);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// MAY_TOP: the pass can hold global stage 0 (bit log_n - 1), which can only
// be its last stage kk = R - 1; passes over lower bits never do, and skip
// the run-time test (and the second copy of every stage it would compile).
template <class A, int R, int NE, bool MAY_TOP = true>
__device__ __forceinline__ void gs_pass(typename A::T (&v)[NE], uint32_t gx0, int lo, int log_n,
                                        const typename A::Twid* __restrict__ tw, const typename A::Prime& c) {
    constexpr int MAXH = NE / 2;
    // 32-bit words: every stage's twiddles are loaded before the first
    // butterfly (a top stage's slot reads tw[1], in bounds and unused), so
    // their latency hides behind the earlier stages; 64-bit twiddles would
    // take too many registers and load stage by stage
    constexpr bool kAhead = sizeof(typename A::T) == 4;
    typename A::Twid ta[kAhead ? R : 1][MAXH];
    if constexpr (kAhead) {
#pragma unroll
        for (int kk = 0; kk < R; ++kk) {
            const int beta = lo + kk;
            const uint32_t base = (1u << (log_n - 1 - beta)) + (gx0 >> (beta + 1));
            const int nh = NE >> (kk + 1);
#pragma unroll
            for (int h = 0; h < MAXH; ++h)
                if (h < nh) ta[kk][h] = tw[base + h];
        }
    }
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
        const int beta = lo + kk;
        if (MAY_TOP && kk == R - 1 && beta == log_n - 1) {  // stage 0: merged N^-1 R scaling
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                if (e & (1 << kk)) continue;
                Bfly<A>::gs_last(v[e], v[e | (1 << kk)], c);
            }
            continue;
        }
        typename A::Twid t[MAXH];
        if constexpr (kAhead) {
#pragma unroll
            for (int h = 0; h < MAXH; ++h) t[h] = ta[kk][h];
        } else {
            const uint32_t base = (1u << (log_n - 1 - beta)) + (gx0 >> (beta + 1));
            const int nh = NE >> (kk + 1);
#pragma unroll
            for (int h = 0; h < MAXH; ++h)
                if (h < nh) t[h] = tw[base + h];
        }
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            if (e & (1 << kk)) continue;
            Bfly<A>::gs(v[e], v[e | (1 << kk)], t[e >> (kk + 1)], c.p, c.p2);
        }
        if constexpr (!kAhead) __builtin_amdgcn_sched_barrier(0);
    }
}

// ---------------------------------------------------------------------------
// inner kernel: one 2^LOGB block of one prime row per (sub-)workgroup
// ---------------------------------------------------------------------------
template <int LOGB>
struct InnerShape {
    static constexpr int TPB = 1 << (LOGB - 4);             // threads per block (16 values each)
    static constexpr int RL = LOGB - 4 * ((LOGB - 1) / 4);  // bits of the last pass (1..4)
    static constexpr int NP = 1 + (LOGB - RL) / 4;          // passes
    static constexpr int BPW = TPB >= 256 ? 1 : 256 / TPB;  // blocks per workgroup
    static constexpr int WG = BPW * TPB;
    static constexpr int PADB = (1 << LOGB) + (1 << (LOGB - 4));  // one pad word per 16
};

// Levels the 30-bit inner block merges (ntt_inner30_block, merge_product30):
// its last MG forward stages, the pointwise product and its first MG inverse
// stages become one product mod X^(2^MG) - C per group of 2^MG adjacent
// positions, which leaves the block's output scaled by 2^-MG; the prime
// constants of variant MG (rows MG k .. MG k + k - 1 of the 30-bit
// PrimeConst32 table) carry 2^MG in the final inverse stage's factors.  Two
// levels where the last pass holds two stages or more, one where it holds
// one; blocks below 2^5 take the one-operand kernel and merge nothing.
__host__ __device__ constexpr uint32_t merge30(uint32_t logb) {
    return logb < 5 ? 0u : (logb - 4 * ((logb - 1) / 4) >= 2 ? 2u : 1u);
}

// Layout of pass m: register e of thread tid holds local index
//   x = base(tid) + delta(e),  delta(e) = e << LO,
// where base fills the other bits.  Pass m < NP - 1 covers the four bits
// [LO, LO + 4) counted down from the top (pass 0: the top four, so x = tid +
// e TPB -- loads and stores in it are lane-contiguous); the last pass covers
// [0, RL), RL <= 4, its registers also spanning bits [RL, 4) side by side.
// LDS rows are padded by one word per 16 (pad(x) = x + (x >> 4)); pad is
// additive over disjoint bit fields, so every per-register LDS and global
// offset below is a compile-time constant added to one per-thread base.
template <int LOGB, int M>
struct PassMap {
    using S = InnerShape<LOGB>;
    static constexpr int LO = M == S::NP - 1 ? 0 : LOGB - 4 * (M + 1);
    static constexpr int R = M == S::NP - 1 ? S::RL : 4;
    static constexpr uint32_t base(uint32_t tid) {
        return (tid & ((1u << LO) - 1)) | ((tid >> LO) << (LO + 4));
    }
    static constexpr uint32_t delta(int e) { return (uint32_t)e << LO; }
    // index bits that one wave's lanes and registers span (base and delta
    // are bit permutations of tid and e)
    static constexpr uint32_t wave_bits() {
        uint32_t m = 0;
        for (int i = 0; i < 6; ++i) m |= base(1u << i);
        for (int i = 0; i < 4; ++i) m |= delta(1 << i);
        return m;
    }
};

// An exchange between two pass layouts is wave-local when every wave holds
// the same set of indices before and after it: the whole block lives in one
// wave, or both layouts give each wave the same spanned bits and the same
// fixed bits.  Such an exchange needs no workgroup barrier: the region a wave
// writes is the one it alone read in the exchange before (whose target
// layout is this one's source), and LDS serves one wave's accesses in order.
template <int LOGB, int FROM, int TO>
constexpr bool exchange_in_wave() {
    using S = InnerShape<LOGB>;
    using F = PassMap<LOGB, FROM>;
    using G = PassMap<LOGB, TO>;
    if (S::TPB <= 64) return true;
    if (F::wave_bits() != G::wave_bits()) return false;
    for (uint32_t w = 0; w < (uint32_t)S::TPB / 64; ++w)
        if (F::base(64 * w) != G::base(64 * w)) return false;
    return true;
}

template <bool IN_WAVE>
__device__ __forceinline__ void exchange_sync() {
    if constexpr (IN_WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

__host__ __device__ constexpr uint32_t lds_pad(uint32_t x) { return x + (x >> 4); }

template <class A, int LOGB, int FROM, int TO>
__device__ __forceinline__ void exchange(typename A::T (&v)[16], typename A::T* lds, uint32_t tid) {
    using F = PassMap<LOGB, FROM>;
    using G = PassMap<LOGB, TO>;
    constexpr bool kWave = exchange_in_wave<LOGB, FROM, TO>();
    typename A::T* w = lds + lds_pad(F::base(tid));
    typename A::T* r = lds + lds_pad(G::base(tid));
    exchange_sync<kWave>();
#pragma unroll
    for (int e = 0; e < 16; ++e) w[lds_pad(F::delta(e))] = v[e];
    exchange_sync<kWave>();
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = r[lds_pad(G::delta(e))];
}

template <class A, int LOGB, int M>
__device__ __forceinline__ void forward_from(typename A::T (&v)[16], typename A::T* lds, uint32_t tid, uint32_t gblk,
                                             int log_n, const typename A::Twid* tw, typename A::T p,
                                             typename A::T p2) {
    using P = PassMap<LOGB, M>;
    if constexpr (M > 0) exchange<A, LOGB, M - 1, M>(v, lds, tid);
    ct_pass<A, P::R, 16>(v, gblk | P::base(tid), P::LO, log_n, tw, p, p2);
    if constexpr (M + 1 < InnerShape<LOGB>::NP) forward_from<A, LOGB, M + 1>(v, lds, tid, gblk, log_n, tw, p, p2);
}

template <class A, int LOGB, int M>
__device__ __forceinline__ void inverse_from(typename A::T (&v)[16], typename A::T* lds, uint32_t tid, uint32_t gblk,
                                             int log_n, const typename A::Twid* tw, const typename A::Prime& c) {
    using P = PassMap<LOGB, M>;
    gs_pass<A, P::R, 16, M == 0>(v, gblk | P::base(tid), P::LO, log_n, tw, c);
    if constexpr (M > 0) {
        exchange<A, LOGB, M, M - 1>(v, lds, tid);
        inverse_from<A, LOGB, M - 1>(v, lds, tid, gblk, log_n, tw, c);
    }
}

// ra, rb: [nprod][k][N] residue rows after the outer forward stages (or
// canonical residues when the block is the whole row).  Output overwrites ra:
// lazy [0, 2p) values awaiting the outer inverse stages, or canonical
// residues of the product when the block is the whole row.
//
// Only one operand lives in registers at a time: a's forward transform is
// parked in a's own slots of ra (in the last pass's layout, so every thread
// re-reads exactly what it wrote), b is transformed, and the pointwise
// product re-reads a.
template <class A, int LOGB>
__global__ __launch_bounds__(InnerShape<LOGB>::WG) void ntt_inner_kernel(
    typename A::T* __restrict__ ra, const typename A::T* __restrict__ rb, const typename A::Twid* __restrict__ twf,
    const typename A::Twid* __restrict__ twi, const typename A::Prime* __restrict__ pcs, uint32_t log_n, uint32_t k,
    uint32_t nprod, uint32_t tw_log) {
    using S = InnerShape<LOGB>;
    using T = typename A::T;
    constexpr int LAST = S::NP - 1;
    __shared__ T lds_all[S::BPW * S::PADB];
    const uint32_t sub = threadIdx.x / S::TPB;
    const uint32_t tid = threadIdx.x % S::TPB;
    T* lds = lds_all + sub * S::PADB;
    const uint32_t N = 1u << log_n;
    const uint32_t nblk_log = log_n - LOGB;
    const uint32_t total = k * nprod << nblk_log;
    uint32_t wid = xcd_remap(blockIdx.x, gridDim.x) * S::BPW + sub;
    if constexpr (S::TPB >= 64) wid = uniform(wid);  // whole waves per block: scalar prime constants
    const bool active = wid < total;
    // an idle sub-block works on block 0 (its loads stay in bounds and
    // unconditional, so none of them needs a branch) for uniform barriers,
    // and stores nothing
    if (!active) wid = 0;
    const uint32_t blk = wid & ((1u << nblk_log) - 1);
    const uint32_t row = wid >> nblk_log;  // = j * nprod + prod
    const uint32_t j = row / nprod, prod = row - j * nprod;
    const size_t off = ((size_t)prod * k + j) * N + ((size_t)blk << LOGB);
    const uint32_t gblk = blk << LOGB;
    const typename A::Prime c = pcs[j];
    const typename A::Twid* tf = twf + ((size_t)j << tw_log);
    const typename A::Twid* ti = twi + ((size_t)j << tw_log);
    const bool final_out = (nblk_log == 0);

    using P0 = PassMap<LOGB, 0>;
    using PL = PassMap<LOGB, LAST>;
    T* ga0 = ra + off + P0::base(tid);  // pass-0 layout (loads / final store)
    T* gal = ra + off + PL::base(tid);  // last-pass layout (parks a)
    const T* gb0 = rb + off + P0::base(tid);
    T v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = ga0[P0::delta(e)];
    forward_from<A, LOGB, 0>(v, lds, tid, gblk, (int)log_n, tf, c.p, c.p2);
    if (active) {
#pragma unroll
        for (int e = 0; e < 16; ++e) gal[PL::delta(e)] = A::sub_if_ge(v[e], c.p2);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = gb0[P0::delta(e)];
    forward_from<A, LOGB, 0>(v, lds, tid, gblk, (int)log_n, tf, c.p, c.p2);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const T x = gal[PL::delta(e)];
        v[e] = A::mont(x, A::sub_if_ge(v[e], c.p2), c.p, c.mont_neg);
    }
    inverse_from<A, LOGB, LAST>(v, lds, tid, gblk, (int)log_n, ti, c);
    if (!active) return;
#pragma unroll
    for (int e = 0; e < 16; ++e) ga0[P0::delta(e)] = final_out ? A::sub_if_ge(v[e], c.p) : v[e];
}

// ---------------------------------------------------------------------------
// 30-bit inner kernel: both operands live (16 + 16 words per thread).  The
// forward transforms of a and b advance stage by stage on the same twiddle
// loads, their exchanges move (a_x, b_x) pairs as 64-bit LDS words (the
// padded layouts stay conflict-free at 8 bytes), the pointwise product needs
// no parking, and the inverse transform runs on the product alone.
// ---------------------------------------------------------------------------
// One pass's twiddles (stage kk: 8 >> kk of them for the 16 registers; a
// pass of R < 4 bits has spare register bits above its own, see ct_pass),
// loaded ahead of the pass: before the LDS exchange that precedes it, so
// their latency hides behind the exchange.  A top stage's slot reads tw[1]
// (in bounds, unused).  Slots of stages a pass does not run stay unloaded
// (and take no registers); the last forward pass also holds the merge
// stage's twiddles in slot MG (merge_product30).
template <int R>
struct PassTw30 {
    Tw32 t[4][8];
};

// Stages kk in [KMIN, R) of a pass, and stage KX >= R as well when KX >= 0
// (the merge stage of a last pass that does not run it).
template <int R, int KMIN = 0, int KX = -1>
__device__ __forceinline__ void load_pass_tw(PassTw30<R>& pt, uint32_t gx0, int lo, int log_n,
                                             const Tw32* __restrict__ tw) {
    static_assert(KX < 4, "four stage slots");
#pragma unroll
    for (int kk = 3; kk >= KMIN; --kk) {
        if (kk >= R && kk != KX) continue;
        const int beta = lo + kk;
        const uint32_t base = (1u << (log_n - 1 - beta)) + (gx0 >> (beta + 1));
#pragma unroll
        for (int h = 0; h < (8 >> kk); ++h) pt.t[kk][h] = tw[base + h];
    }
}

template <class A, int R, int KMIN = 0>
__device__ __forceinline__ void ct_pass_pair(typename A::T (&va)[16], typename A::T (&vb)[16], const PassTw30<R>& pt,
                                             typename A::T p, typename A::T p2) {
#pragma unroll
    for (int kk = R - 1; kk >= KMIN; --kk) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            if (e & (1 << kk)) continue;
            Bfly<A>::ct(va[e], va[e | (1 << kk)], pt.t[kk][e >> (kk + 1)], p, p2);
            Bfly<A>::ct(vb[e], vb[e | (1 << kk)], pt.t[kk][e >> (kk + 1)], p, p2);
        }
    }
}

// The last MG forward stages of a and b (stages 0 .. MG - 1 of the last pass,
// whose registers e hold positions x0 + e), the pointwise product and the
// first MG inverse stages, as one product mod X^G - C (G = 2^MG) per group of
// G adjacent registers: the stage-MG butterflies leave each group as the
// operand mod X^G - C, coefficients in register order, with C = w^2 for the
// group's stage-(MG - 1) twiddle w, i.e. C = +/- t for its stage-MG twiddle t
// (minus where bit MG of the position is set).  In: a, b in [0, 4p).  Out
// (va): coefficient k of the product times 2^-32, in [0, 2p) -- 2^-MG times
// what the separate stages give (the variant-MG prime constants restore it).
//   a -> [0, p); b -> [0, 2p); C b_j by Shoup on the stored negated twiddle
//   (+t b = -nv, -t b = nv + 2p, both in [0, 2p]); each coefficient
//   T = sum_i a_i (b_(k-i) or C b_(k-i+G)) is one 64-bit sum below
//   2 G p^2 <= 8 p^2, reduced as (T + m p) / 2^32 (m = T (-p^-1) mod 2^32)
//   < 2 G p^2 / 2^32 + p: below 2p for G = 2 and below 3p for G = 4, which
//   takes one more conditional subtraction (4p < 2^32).
template <int MG>
__device__ __forceinline__ void merge_product30(uint32_t (&va)[16], const uint32_t (&vb)[16], const Tw32 (&mt)[8],
                                                const PrimeConst32& c) {
    static_assert(MG == 1 || MG == 2, "one or two merged levels");
    constexpr int G = 1 << MG;
    const uint32_t p = c.p, p2 = c.p2, mneg = c.mont_neg;
#pragma unroll
    for (int g0 = 0; g0 < 16; g0 += G) {
        const Tw32 w = mt[g0 >> (MG + 1)];
        uint32_t a[G], b[G], bc[G];
#pragma unroll
        for (int i = 0; i < G; ++i) {
            a[i] = W30::sub_if_ge(W30::sub_if_ge(va[g0 + i], p2), p);
            b[i] = W30::sub_if_ge(vb[g0 + i], p2);
        }
#pragma unroll
        for (int j = 1; j < G; ++j) {
            const uint32_t nv = __umulhi(vb[g0 + j], w.ws) * p + vb[g0 + j] * w.w;
            bc[j] = ((g0 >> MG) & 1) ? nv + p2 : 0u - nv;
        }
#pragma unroll
        for (int kx = 0; kx < G; ++kx) {
            uint64_t t = 0;
#pragma unroll
            for (int i = 0; i < G; ++i) {
                const int jj = kx - i;
                t += (uint64_t)a[i] * (jj >= 0 ? b[jj] : bc[jj + G]);
            }
            const uint32_t m = (uint32_t)t * mneg;
            const uint32_t r = (uint32_t)((t + (uint64_t)m * p) >> 32);
            va[g0 + kx] = MG == 1 ? r : W30::sub_if_ge(r, p2);
        }
    }
}

// The inverse pass of 30-bit words on preloaded twiddles (MAY_TOP: as
// gs_pass), from stage KMIN (the first inverse pass skips the merged ones).
template <int R, int NE, bool MAY_TOP, int KMIN = 0>
__device__ __forceinline__ void gs_pass_tw(uint32_t (&v)[NE], const PassTw30<R>& pt, int lo, int log_n,
                                           const PrimeConst32& c) {
#pragma unroll
    for (int kk = KMIN; kk < R; ++kk) {
        const int beta = lo + kk;
        if (MAY_TOP && kk == R - 1 && beta == log_n - 1) {  // stage 0: merged N^-1 R scaling
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                if (e & (1 << kk)) continue;
                Bfly<W30>::gs_last(v[e], v[e | (1 << kk)], c);
            }
            continue;
        }
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            if (e & (1 << kk)) continue;
            Bfly<W30>::gs(v[e], v[e | (1 << kk)], pt.t[kk][e >> (kk + 1)], c.p, c.p2);
        }
    }
}

template <int LOGB, int FROM, int TO>
__device__ __forceinline__ void exchange_pair(uint32_t (&va)[16], uint32_t (&vb)[16], uint2* lds, uint32_t tid) {
    using F = PassMap<LOGB, FROM>;
    using G = PassMap<LOGB, TO>;
    constexpr bool kWave = exchange_in_wave<LOGB, FROM, TO>();
    uint2* w = lds + lds_pad(F::base(tid));
    uint2* r = lds + lds_pad(G::base(tid));
    exchange_sync<kWave>();
#pragma unroll
    for (int e = 0; e < 16; ++e) w[lds_pad(F::delta(e))] = make_uint2(va[e], vb[e]);
    exchange_sync<kWave>();
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const uint2 x = r[lds_pad(G::delta(e))];
        va[e] = x.x;
        vb[e] = x.y;
    }
}

// PF: each pass's twiddles are loaded before the exchange that precedes it
// (the caller loads pt); otherwise at the start of the pass (pt is loaded
// here).  Prefetching keeps one more pass's twiddles live across the
// exchange, which only some block sizes have the registers for.  The last
// pass stops above the merged stages and ends in merge_product30: va then
// holds the merged product (2^-MG scaled).
template <int LOGB, int M>
struct FwdPass30 {
    static constexpr bool kLast = M + 1 == InnerShape<LOGB>::NP;
    static constexpr int MG = (int)merge30(LOGB);
    static constexpr int KMIN = kLast ? MG : 0;  // lowest stage the pass runs
    static constexpr int KX = kLast ? MG : -1;   // merge twiddles in slot MG
    using P = PassMap<LOGB, M>;
    static __device__ __forceinline__ void load(PassTw30<P::R>& pt, uint32_t gblk, uint32_t tid, int log_n,
                                                const Tw32* tw) {
        load_pass_tw<P::R, KMIN, KX>(pt, gblk | P::base(tid), P::LO, log_n, tw);
    }
};

template <int LOGB, int M, bool PF>
__device__ __forceinline__ void forward_pair_from(uint32_t (&va)[16], uint32_t (&vb)[16], uint2* lds, uint32_t tid,
                                                  uint32_t gblk, int log_n, const Tw32* tw, const PrimeConst32& c,
                                                  PassTw30<PassMap<LOGB, M>::R>& pt) {
    using F = FwdPass30<LOGB, M>;
    using P = typename F::P;
    if constexpr (M > 0) exchange_pair<LOGB, M - 1, M>(va, vb, lds, tid);
    if constexpr (!PF) F::load(pt, gblk, tid, log_n, tw);
    ct_pass_pair<W30, P::R, F::KMIN>(va, vb, pt, c.p, c.p2);
    if constexpr (!F::kLast) {
        using G = FwdPass30<LOGB, M + 1>;
        PassTw30<G::P::R> nt;
        if constexpr (PF) G::load(nt, gblk, tid, log_n, tw);
        forward_pair_from<LOGB, M + 1, PF>(va, vb, lds, tid, gblk, log_n, tw, c, nt);
    } else {
        merge_product30<F::MG>(va, vb, pt.t[F::MG], c);
    }
}

template <int LOGB, int M, bool PF>
__device__ __forceinline__ void inverse30_from(uint32_t (&v)[16], uint32_t* lds, uint32_t tid, uint32_t gblk,
                                               int log_n, const Tw32* tw, const PrimeConst32& c,
                                               PassTw30<PassMap<LOGB, M>::R>& pt) {
    using P = PassMap<LOGB, M>;
    // the first inverse pass (the last pass's layout) starts above the merged stages
    constexpr int KMIN = M + 1 == InnerShape<LOGB>::NP ? (int)merge30(LOGB) : 0;
    if constexpr (!PF) load_pass_tw<P::R, KMIN>(pt, gblk | P::base(tid), P::LO, log_n, tw);
    gs_pass_tw<P::R, 16, M == 0, KMIN>(v, pt, P::LO, log_n, c);
    if constexpr (M > 0) {
        using Q = PassMap<LOGB, M - 1>;
        PassTw30<Q::R> nt;
        if constexpr (PF) load_pass_tw<Q::R>(nt, gblk | Q::base(tid), Q::LO, log_n, tw);
        exchange<W30, LOGB, M, M - 1>(v, lds, tid);
        inverse30_from<LOGB, M - 1, PF>(v, lds, tid, gblk, log_n, tw, c, nt);
    }
}

// One 2^LOGB block of one prime row: ra_blk / rb_blk point at the block in
// a's and b's rows (the product overwrites ra_blk), gblk is the block's first
// row index, tf / ti the prime's twiddle rows, c the prime's constants of
// variant merge30(LOGB).  tid: the thread's index in its (sub-)block; lds2:
// the sub-block's padded region.  Every thread of the workgroup calls this
// (barriers inside); an inactive sub-block computes and stores nothing it is
// not given to store.
template <int LOGB>
__device__ __forceinline__ void ntt_inner30_block(uint2* lds2, uint32_t tid, uint32_t* __restrict__ ra_blk,
                                                  const uint32_t* __restrict__ rb_blk, uint32_t gblk,
                                                  const PrimeConst32& c, const Tw32* tf, const Tw32* ti,
                                                  uint32_t log_n, bool active, bool final_out) {
    static_assert(LOGB >= 5, "blocks of 2^5 and up (merge30)");
    using S = InnerShape<LOGB>;
    using T = uint32_t;
    constexpr int LAST = S::NP - 1;
    T* lds1 = reinterpret_cast<T*>(lds2);  // the inverse transform reuses the region for one operand
    using P0 = PassMap<LOGB, 0>;  // lane-contiguous: x = tid + e TPB
    T* ga0 = ra_blk + P0::base(tid);
    const T* gb0 = rb_blk + P0::base(tid);
    T va[16], vb[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        va[e] = ga0[P0::delta(e)];
        vb[e] = gb0[P0::delta(e)];
    }
    // every pass's twiddles load before the exchange that precedes it from
    // blocks of 2^10 up (it measured faster or the same there and fits in
    // 128 VGPRs); smaller blocks load at the start of each pass
    constexpr bool PF = LOGB >= 10;
    PassTw30<P0::R> tw0;
    if constexpr (PF) FwdPass30<LOGB, 0>::load(tw0, gblk, tid, (int)log_n, tf);
    // forward stages down to the merged ones, then the merged product
    forward_pair_from<LOGB, 0, PF>(va, vb, lds2, tid, gblk, (int)log_n, tf, c, tw0);
    using PL = PassMap<LOGB, LAST>;
    PassTw30<PL::R> twl;
    if constexpr (PF) load_pass_tw<PL::R, (int)merge30(LOGB)>(twl, gblk | PL::base(tid), PL::LO, (int)log_n, ti);
    __syncthreads();  // the paired region is reused below with the one-operand layout
    inverse30_from<LOGB, LAST, PF>(va, lds1, tid, gblk, (int)log_n, ti, c, twl);
    if (!active) return;
#pragma unroll
    for (int e = 0; e < 16; ++e) ga0[P0::delta(e)] = final_out ? W30::sub_if_ge(va[e], c.p) : va[e];
}

template <int LOGB>
__global__ __launch_bounds__(InnerShape<LOGB>::WG) void ntt_inner30_kernel(
    uint32_t* __restrict__ ra, const uint32_t* __restrict__ rb, const Tw32* __restrict__ twf,
    const Tw32* __restrict__ twi, const PrimeConst32* __restrict__ pcs, uint32_t log_n, uint32_t k, uint32_t nprod,
    uint32_t tw_log) {
    using S = InnerShape<LOGB>;
    __shared__ uint2 lds_all[S::BPW * S::PADB];
    const uint32_t sub = threadIdx.x / S::TPB;
    const uint32_t tid = threadIdx.x % S::TPB;
    const uint32_t N = 1u << log_n;
    const uint32_t nblk_log = log_n - LOGB;
    const uint32_t total = k * nprod << nblk_log;
    uint32_t wid = xcd_remap(blockIdx.x, gridDim.x) * S::BPW + sub;
    if constexpr (S::TPB >= 64) wid = uniform(wid);
    const bool active = wid < total;
    // an idle sub-block works on block 0 (its loads stay in bounds and
    // unconditional, so none of them needs a branch) for uniform barriers,
    // and stores nothing
    if (!active) wid = 0;
    const uint32_t blk = wid & ((1u << nblk_log) - 1);
    const uint32_t row = wid >> nblk_log;  // = j * nprod + prod
    const uint32_t j = row / nprod, prod = row - j * nprod;
    const size_t off = ((size_t)prod * k + j) * N + ((size_t)blk << LOGB);
    const PrimeConst32 c = pcs[j];
    ntt_inner30_block<LOGB>(lds_all + sub * S::PADB, tid, ra + off, rb + off, blk << LOGB, c,
                            twf + ((size_t)j << tw_log), twi + ((size_t)j << tw_log), log_n, active,
                            nblk_log == 0);
}

// ---------------------------------------------------------------------------
// outer kernels: the first S1 forward stages / last S1 inverse stages.
// Thread = one column x = col + (e << b), e in [0, 2^S1), b = logN - S1.
// ---------------------------------------------------------------------------
template <class A, int S1>
__global__ __launch_bounds__(256) void ntt_outer_fwd_kernel(typename A::T* __restrict__ ra,
                                                            typename A::T* __restrict__ rb,
                                                            const typename A::Twid* __restrict__ twf,
                                                            const typename A::Prime* __restrict__ pcs,
                                                            uint32_t log_n, uint32_t k, uint32_t nprod,
                                                            uint32_t tw_log) {
    using T = typename A::T;
    constexpr int NE = 1 << S1;
    const uint32_t b = log_n - S1;
    const uint32_t cols = 1u << b;
    const uint32_t wgs_per_row = (cols + 255) / 256;
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t row = g / wgs_per_row;  // ((j * 2 + which) * nprod + prod)
    const uint32_t col = (g - row * wgs_per_row) * 256 + threadIdx.x;
    if (col >= cols) return;
    const uint32_t prod = row % nprod;
    const uint32_t jw = row / nprod;
    const uint32_t which = jw & 1u, j = jw >> 1;
    T* base = (which ? rb : ra) + ((size_t)prod * k + j) * ((size_t)1 << log_n) + col;
    const typename A::Prime c = pcs[j];
    T v[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) v[e] = base[(size_t)e << b];
    ct_pass<A, S1, NE>(v, col, (int)b, (int)log_n, twf + ((size_t)j << tw_log), c.p, c.p2);
#pragma unroll
    for (int e = 0; e < NE; ++e) base[(size_t)e << b] = v[e];
}

template <class A, int S1>
__global__ __launch_bounds__(256) void ntt_outer_inv_kernel(typename A::T* __restrict__ ra,
                                                            const typename A::Twid* __restrict__ twi,
                                                            const typename A::Prime* __restrict__ pcs,
                                                            uint32_t log_n, uint32_t k, uint32_t nprod,
                                                            uint32_t tw_log) {
    using T = typename A::T;
    constexpr int NE = 1 << S1;
    const uint32_t b = log_n - S1;
    const uint32_t cols = 1u << b;
    const uint32_t wgs_per_row = (cols + 255) / 256;
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t row = g / wgs_per_row;  // (j * nprod + prod)
    const uint32_t col = (g - row * wgs_per_row) * 256 + threadIdx.x;
    if (col >= cols) return;
    const uint32_t j = row / nprod, prod = row - j * nprod;
    T* base = ra + ((size_t)prod * k + j) * ((size_t)1 << log_n) + col;
    const typename A::Prime c = pcs[j];
    T v[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) v[e] = base[(size_t)e << b];
    gs_pass<A, S1, NE>(v, col, (int)b, (int)log_n, twi + ((size_t)j << tw_log), c);
#pragma unroll
    for (int e = 0; e < NE; ++e) base[(size_t)e << b] = A::sub_if_ge(v[e], c.p);
}

// ---------------------------------------------------------------------------
// tiny rows (N < 16): one thread per (product, prime), plain loops.
// ---------------------------------------------------------------------------
template <class A>
__global__ __launch_bounds__(64) void ntt_tiny_kernel(typename A::T* __restrict__ ra,
                                                      const typename A::T* __restrict__ rb,
                                                      const typename A::Twid* __restrict__ twf,
                                                      const typename A::Twid* __restrict__ twi,
                                                      const typename A::Prime* __restrict__ pcs, uint32_t log_n,
                                                      uint32_t k, uint32_t nprod, uint32_t tw_log) {
    using T = typename A::T;
    const uint32_t t = blockIdx.x * 64 + threadIdx.x;
    if (t >= k * nprod) return;
    const uint32_t j = t / nprod, prod = t - j * nprod;
    const uint32_t N = 1u << log_n;
    const size_t off = ((size_t)prod * k + j) * N;
    const typename A::Prime c = pcs[j];
    const typename A::Twid* tf = twf + ((size_t)j << tw_log);
    const typename A::Twid* ti = twi + ((size_t)j << tw_log);
    T x[8], y[8];
    for (uint32_t i = 0; i < N; ++i) {
        x[i] = ra[off + i];
        y[i] = rb[off + i];
    }
    for (int s = 0; s < (int)log_n; ++s) {
        const int beta = (int)log_n - 1 - s;
        for (uint32_t i = 0; i < N; ++i) {
            if (i & (1u << beta)) continue;
            const typename A::Twid w = tf[(1u << s) + (i >> (beta + 1))];
            Bfly<A>::ct(x[i], x[i | (1u << beta)], w, c.p, c.p2);
            Bfly<A>::ct(y[i], y[i | (1u << beta)], w, c.p, c.p2);
        }
    }
    for (uint32_t i = 0; i < N; ++i)
        x[i] = A::mont(A::sub_if_ge(x[i], c.p2), A::sub_if_ge(y[i], c.p2), c.p, c.mont_neg);
    for (int beta = 0; beta < (int)log_n; ++beta) {
        const int s = (int)log_n - 1 - beta;
        for (uint32_t i = 0; i < N; ++i) {
            if (i & (1u << beta)) continue;
            if (s == 0) Bfly<A>::gs_last(x[i], x[i | (1u << beta)], c);
            else Bfly<A>::gs(x[i], x[i | (1u << beta)], ti[(1u << s) + (i >> (beta + 1))], c.p, c.p2);
        }
    }
    for (uint32_t i = 0; i < N; ++i) ra[off + i] = A::sub_if_ge(x[i], c.p);
}

// ===========================================================================
// CRT: CRT-weighted residues -> L limbs mod q.
//   v_j   = r_j (M/p_j)^-1 mod p_j                   (emitted by the inverse
//                                                     NTT, canonical)
//   acc   = sum_j v_j ((M/p_j) mod q)                (LT + 3 limbs)
//   frac  = sum_j floor(v_j 2^64 / p_j), each estimated from below by at
//           most one unit (128-bit fixed point)
//   K     = floor(frac / 2^64) + [frac mod 2^64 >= 2^63]
//   c     = (acc + K ((-M) mod q)) mod q
// K counts the multiples of M in sum_j v_j M/p_j plus the centred-lift
// correction, exactly as solve.cu:779-834 subtracts sum_hi (M mod q) and,
// when the fraction is >= 1/2, M mod q once more; folding both into one
// non-negative addition leaves a single Barrett reduction.
// ===========================================================================
template <int NA>
__device__ __forceinline__ void mac_limbs(uint32_t (&acc)[NA], const uint32_t* __restrict__ m, int mlen_ct,
                                          uint32_t f, int off) {
    // carry-save: every product f m_i is independent; its low word joins
    // limb off + i and its high word limb off + i + 1 through two 32-bit
    // add-with-carry chains (one v_addc_co_u32 each per limb), the carry
    // out of the top limb is dropped
    uint32_t ca = 0, cb = 0, hi_prev = 0;
#pragma unroll
    for (int l = 0; l < NA; ++l) {
        if (l < off) continue;
        uint32_t lo = 0, hi = 0;
        if (l - off < mlen_ct) {
            const uint64_t t = (uint64_t)f * m[l - off];
            lo = (uint32_t)t;
            hi = (uint32_t)(t >> 32);
        }
        uint32_t na, nb;
        const uint32_t a = __builtin_addc(acc[l], lo, ca, &na);
        acc[l] = __builtin_addc(a, hi_prev, cb, &nb);
        ca = na;
        cb = nb;
        hi_prev = hi;
    }
}

// acc mod q for acc < 2^(bits(q) + 71), by a Barrett quotient estimate
// qhat = floor((acc >> (bits(q)-1)) * mu / 2^72), mu = floor(2^(bits(q)-1+72)/q),
// at most 2 below the true quotient; then <= 2 subtractions of q.
template <int LT>
__device__ __forceinline__ void barrett_mod_q(uint32_t (&acc)[LT + 3], const QConst* __restrict__ qc) {
    constexpr int NA = LT + 3;
    const uint32_t sh = qc->q_bits - 1;
    const uint32_t w = sh >> 5, bs = sh & 31u;
    // t = acc >> sh, three limbs (selected without dynamic register indexing)
    uint32_t t[3] = {0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        uint32_t lo = 0, hi = 0;
#pragma unroll
        for (int l = 0; l < NA; ++l) {
            if ((uint32_t)l == w + i) lo = acc[l];
            if ((uint32_t)l == w + i + 1) hi = acc[l];
        }
        t[i] = bs ? ((lo >> bs) | (hi << (32 - bs))) : lo;
    }
    uint32_t pr[6] = {0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        uint64_t carry = 0;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const uint64_t x = (uint64_t)t[i] * qc->mu[jj] + pr[i + jj] + carry;
            pr[i + jj] = (uint32_t)x;
            carry = x >> 32;
        }
        pr[i + 3] = (uint32_t)carry;
    }
    uint32_t qh[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) qh[i] = (pr[2 + i] >> 8) | (pr[3 + i] << 24);
    // acc -= qhat * q, one multiply-subtract row per qhat limb
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        uint32_t borrow = 0;
#pragma unroll
        for (int l = 0; l < NA; ++l) {
            if (l < i) continue;
            const uint64_t prod = (l - i < LT ? (uint64_t)qh[i] * qc->q[l - i] : 0ull) + borrow;
            const uint32_t lo = (uint32_t)prod;
            borrow = (uint32_t)(prod >> 32) + (acc[l] < lo);
            acc[l] -= lo;
        }
    }
#pragma unroll
    for (int fix = 0; fix < 2; ++fix) {
        // acc >= q ?  compare from the top, then subtract under the flag
        int ge = 1;
        bool decided = false;
#pragma unroll
        for (int l = NA - 1; l >= 0; --l) {
            const uint32_t ql = l < LT ? qc->q[l] : 0u;
            const bool gt = acc[l] > ql, lt = acc[l] < ql;
            if (!decided && (gt || lt)) ge = gt ? 1 : 0;
            decided = decided || gt || lt;
        }
        const uint32_t mask = 0u - (uint32_t)ge;
        uint32_t br = 0;
#pragma unroll
        for (int l = 0; l < NA; ++l) {
            const uint64_t d = (uint64_t)acc[l] - ((l < LT ? qc->q[l] : 0u) & mask) - br;
            acc[l] = (uint32_t)d;
            br = (uint32_t)(d >> 63);
        }
    }
}

// acc + K ((-M) mod q) reduced mod q, in place (acc[0 .. LT) on return).
template <int LT>
__device__ __forceinline__ void crt_finish(uint32_t (&acc)[LT + 3], uint64_t f_lo, uint64_t f_hi,
                                           const QConst* __restrict__ qc) {
    const uint32_t K = (uint32_t)f_hi + (uint32_t)(f_lo >> 63);
    mac_limbs<LT + 3>(acc, qc->neg_m, LT, K, 0);
    barrett_mod_q<LT>(acc, qc);
}

// A wave's results leave through LDS: lane i holds the limbs of coefficient
// base + i (count lanes valid), and dst = out + base L receives count * L
// consecutive words as coalesced stores instead of one L-word stride per lane.
template <int NR>
__device__ __forceinline__ void wave_store_limbs(const uint32_t* r, uint32_t* __restrict__ lds_wave,
                                                 uint32_t* __restrict__ dst, uint32_t lane, uint32_t L,
                                                 uint32_t count) {
    if (lane < count) {
#pragma unroll
        for (int l = 0; l < NR; ++l)
            if ((uint32_t)l < L) lds_wave[lane * L + l] = r[l];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t words = count * L;
    for (uint32_t w = lane; w < words; w += 64u) dst[w] = lds_wave[w];
}

template <int LT>
__global__ __launch_bounds__(256) void crt62_kernel(const uint64_t* __restrict__ res, uint32_t* __restrict__ out,
                                                    const CrtConst* __restrict__ cc, const uint32_t* __restrict__ miq,
                                                    const QConst* __restrict__ qc, const PrimeConst* __restrict__ pcs,
                                                    uint32_t log_n, uint32_t L, uint32_t k, uint32_t nprod) {
    constexpr int NA = LT + 3;
    __shared__ uint32_t lds_out[256 * LT];
    const uint32_t N = 1u << log_n;
    const uint32_t idx = blockIdx.x * 256 + threadIdx.x;
    const uint32_t total = nprod << log_n;
    const uint32_t lane = threadIdx.x & 63u, wbase = idx - lane;
    if (wbase >= total) return;  // wave-uniform
    const uint32_t count = total - wbase < 64u ? total - wbase : 64u;
    const uint32_t ic = idx < total ? idx : wbase;  // idle lanes recompute a live coefficient, store nothing
    const uint32_t prod = ic >> log_n, i = ic & (N - 1);
    const uint64_t* r = res + (size_t)prod * k * N + i;
    uint32_t acc[NA];
#pragma unroll
    for (int l = 0; l < NA; ++l) acc[l] = 0;
    uint64_t f_lo = 0, f_hi = 0;
    for (uint32_t j0 = 0; j0 < k; j0 += 4) {
        uint64_t vv[4];  // four loads in flight per trip
#pragma unroll
        for (int e = 0; e < 4; ++e) vv[e] = r[(size_t)min(j0 + e, k - 1) * N];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t j = j0 + e;
            if (j >= k) break;
            const CrtConst cj = cc[j];
            const uint64_t v = vv[e];
            const uint64_t f = v * cj.mag_hi + __umul64hi(v, cj.mag_lo);
            f_lo += f;
            f_hi += (f_lo < f);
            const uint32_t* m = miq + (size_t)j * kMaxLimbs;
            mac_limbs<NA>(acc, m, LT, (uint32_t)v, 0);
            mac_limbs<NA>(acc, m, LT, (uint32_t)(v >> 32), 1);
        }
    }
    crt_finish<LT>(acc, f_lo, f_hi, qc);
    wave_store_limbs<LT>(acc, lds_out + (threadIdx.x - lane) * LT, out + (size_t)wbase * L, lane, L, count);
}

// 30-bit primes: v_j < 2^30, so v_j m_l < 2^62 and three products fit a
// 64-bit column; the columns are carried into 32-bit limbs every third prime.
template <int LT>
__global__ __launch_bounds__(256) void crt30_kernel(const uint32_t* __restrict__ res, uint32_t* __restrict__ out,
                                                    const CrtConst32* __restrict__ cc,
                                                    const uint32_t* __restrict__ miq, const QConst* __restrict__ qc,
                                                    const PrimeConst32* __restrict__ pcs, uint32_t log_n, uint32_t L,
                                                    uint32_t k, uint32_t nprod) {
    constexpr int NA = LT + 3;
    __shared__ uint32_t lds_out[256 * LT];
    const uint32_t N = 1u << log_n;
    const uint32_t idx = blockIdx.x * 256 + threadIdx.x;
    const uint32_t total = nprod << log_n;
    const uint32_t lane = threadIdx.x & 63u, wbase = idx - lane;
    if (wbase >= total) return;  // wave-uniform
    const uint32_t count = total - wbase < 64u ? total - wbase : 64u;
    const uint32_t ic = idx < total ? idx : wbase;  // idle lanes recompute a live coefficient, store nothing
    const uint32_t prod = ic >> log_n, i = ic & (N - 1);
    const uint32_t* r = res + (size_t)prod * k * N + i;
    uint64_t col[LT + 1];
#pragma unroll
    for (int l = 0; l <= LT; ++l) col[l] = 0;
    uint64_t f_lo = 0, f_hi = 0;
    uint32_t pending = 0;
    for (uint32_t j0 = 0; j0 < k; j0 += 4) {
        uint32_t vv[4];  // four loads in flight per trip
#pragma unroll
        for (int e = 0; e < 4; ++e) vv[e] = r[(size_t)min(j0 + e, k - 1) * N];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t j = j0 + e;
            if (j >= k) break;
            const CrtConst32 cj = cc[j];
            const uint32_t v = vv[e];
            // floor(v g / 2^32), g = floor(2^96 / p) = g_hi 2^64 + g_lo
            const uint64_t f = ((uint64_t)(v * cj.g_hi) << 32) + (uint64_t)v * (uint32_t)(cj.g_lo >> 32) +
                               __umulhi(v, (uint32_t)cj.g_lo);
            f_lo += f;
            f_hi += (f_lo < f);
            const uint32_t* m = miq + (size_t)j * kMaxLimbs;
#pragma unroll
            for (int l = 0; l < LT; ++l) col[l] += (uint64_t)v * m[l];
            if (++pending == 3) {  // carry: col[l] < 3 2^62 + 2^33 < 2^64
                pending = 0;
#pragma unroll
                for (int l = 0; l < LT; ++l) {
                    col[l + 1] += col[l] >> 32;
                    col[l] = (uint32_t)col[l];
                }
            }
        }
    }
    uint32_t acc[NA];
    uint64_t carry = 0;
#pragma unroll
    for (int l = 0; l < NA; ++l) {
        const uint64_t t = (l <= LT ? col[l] : 0ull) + carry;
        acc[l] = (uint32_t)t;
        carry = t >> 32;
    }
    crt_finish<LT>(acc, f_lo, f_hi, qc);
    wave_store_limbs<LT>(acc, lds_out + (threadIdx.x - lane) * LT, out + (size_t)wbase * L, lane, L, count);
}

// x -> x 2^-32 mod q for x < (2^32 - 1) q (NA limbs), one word-level
// Montgomery step and one conditional subtraction, in place: the LQ limbs of
// the result (zero beyond L) end in x[1 .. 1 + LQ).  (x + m q) / 2^32 with
// m < 2^32 is below 2q, so one subtraction finishes it.  The subtraction runs
// as a borrow-only pass (r >= q ?) and a masked pass, so no second limb array
// is live.
template <int LQ, int NA>
__device__ __forceinline__ void redc32(uint32_t (&x)[NA], const QConst* __restrict__ qc, uint32_t q_inv) {
    static_assert(NA >= LQ + 2, "room for the step and its carry");
    mac_limbs<NA>(x, qc->q, LQ, x[0] * q_inv, 0);  // x[0] becomes 0
    uint32_t br = 0;  // r = x[1 ..] < 2q: borrow of r - q
#pragma unroll
    for (int l = 0; l <= LQ; ++l) {
        uint32_t nb;
        (void)__builtin_subc(x[1 + l], l < LQ ? qc->q[l] : 0u, br, &nb);
        br = nb;
    }
    const bool ge = br == 0;  // subtract q iff r >= q
    br = 0;
#pragma unroll
    for (int l = 0; l < LQ; ++l) {
        uint32_t nb;
        x[1 + l] = __builtin_subc(x[1 + l], ge ? qc->q[l] : 0u, br, &nb);
        br = nb;
    }
}

// 30-bit primes on the matrix cores.  The weighted sum is an int8
// contraction over kappa = 4 j + u (byte u of v_j): table rows hold the bytes
// of 2^(8u) ((M/p_j) mod q) 2^32 mod q, of 2^(8u) floor(2^56 / p_j) and a row
// of ones (layout and offsets: kCrtRowStride and nc_tables.cpp
// build_crt_mfma).  A wave takes 64 coefficients as two 32-coefficient tiles;
// the B operand of lane (c, h) is simply the four words v_j of primes
// 8 kc + 4 h + e of its tile's coefficient c (xor 0x80808080 makes them
// signed bytes), and the output gives lane half h the pseudo-limbs
// lambda = 8 rt + 2 m + h.  One v_permlane32_swap per word hands lane h all
// pseudo-limbs of tile h's coefficient c, and every lane finishes one
// coefficient: carry the 64-bit limb sums, round the fraction to K, add
// K ((-M) mod q) 2^32, and a Montgomery step removes the 2^32: the sum is a
// byte-weighted sum sum_kappa b_kappa m_kappa of at most 256 terms b < 2^8,
// m < q, so with the K term it stays below (2^16 + 2^7) q, well inside the
// single step's range.
// KCT > 0 fixes the number of contraction chunks at compile time (kc_n ==
// KCT); KCT == 0 takes kc_n <= kCrtMaxKC at run time.  LT > 0 fixes L.
// A wave's 64 results (64 L contiguous words of out) leave through its own
// LDS region as 16-byte stores.
//
// INV == 2 also runs the last two inverse NTT stages (what
// ntt_outer_inv_kernel<W30, 2> does between the inner kernel and the CRT):
// a wave then takes positions x0 + i (i < 16) with their partners N/4, N/2
// and 3N/4 away -- tile 0 column c holds x0 + (c & 15) + (c >> 4) N/2,
// tile 1 the same + N/4 -- so the stage of span N/4 pairs a lane's own two
// tiles and the stage of span N/2 (N^-1 R y_j merged) pairs lanes c and
// c ^ 16, and the results leave as four runs of 16 coefficients.  The
// residues enter the contraction lazy, in [0, 2p) rather than canonical:
// the sum x = sum_j v_j (M / p_j) - K M is the same integer for any
// representatives, and the bounds hold for every v_j < 2^31 -- F < 2^63,
// K = round(F / 2^56) <= 127, the limb sum below 2^(32 L + 17) and x < 2^17 q
// for the Montgomery finish -- which spares one conditional subtraction per
// residue (6-7 % of the kernel's time).
// One prime's constants for the fused inverse stages as one half of a
// column group sees them (entry 2 r + u, u = 1 for lanes c with c & 16):
// the span-N/4 twiddle psi^-bitrev(2 + u) and the stage-0 factors, N^-1 R y_j
// for the sum and that times psi^-bitrev(1) for the difference (each with its
// Shoup quotient), so that no lane selects between constants.
struct CrtInvLds {
    uint32_t p, w, ws, nr, nr_s, wr, wr_s, pad;
};

// LDS of the matrix-core CRT: the int8 table (row pitch PITCH bytes), the
// output staging of the 4 waves (STAGE lanes of LW words each: 64, or 32 when
// a wave stages one lane half at a time) and, for the fused inverse stages,
// NINV per-lane-half prime constants.
template <int RT, int PITCH, int LW, int STAGE, int NINV>
struct CrtLds {
    int8_t tab[RT * 32 * PITCH];
    __attribute__((aligned(16))) uint32_t out[4 * STAGE * LW];
    __attribute__((aligned(16))) CrtInvLds inv[NINV];
};

template <int RT, int KCT, int LT, int INV, int PITCH = kCrtRowStride, int STAGE = 64>
using CrtLdsOf = CrtLds<RT, PITCH, (LT > 0 ? LT : kMaxLimbs), STAGE,
                        (INV ? 2 * 8 * (KCT > 0 ? KCT : kCrtMaxKC) : 1)>;

// Loads the CRT's tables into s (every thread of a 256-thread workgroup
// takes part; the caller synchronises before the first group).
template <int RT, int KCT, int LT, int INV, int PITCH, class Lds>
__device__ __forceinline__ void crt30_mfma_tables(Lds& s, const int8_t* __restrict__ tab, uint32_t k, uint32_t kc_n,
                                                  const PrimeConst32* __restrict__ pcs, const Tw32* __restrict__ twi,
                                                  uint32_t tw_log) {
    constexpr int KCM = KCT > 0 ? KCT : kCrtMaxKC;
    if (KCT > 0) kc_n = KCT;
    const uint32_t kbytes = 32 * kc_n;
    const uint32_t chunks = kbytes / 16;
    // the table is row-major with rows of kbytes: piece w is at byte 16 w
    copy16_to_lds<(RT * 32 * KCM * 2 + 255) / 256>(
        RT * 32 * chunks, [&](uint32_t w) { return reinterpret_cast<const nc_v4i*>(tab + (size_t)w * 16); },
        [&](uint32_t w) {
            const uint32_t row = w / chunks, ch = w - row * chunks;
            return reinterpret_cast<nc_v4i*>(s.tab + row * PITCH + ch * 16);
        });
    // entries 8 kc + 4 h + e for every chunk; rows >= k repeat row k - 1
    if constexpr (INV > 0) {
        for (uint32_t r2 = threadIdx.x; r2 < 16u * KCM; r2 += 256) {
            const uint32_t r = r2 >> 1, u = r2 & 1u;
            const uint32_t j = r < k ? r : k - 1;
            const PrimeConst32 c = pcs[j];
            const Tw32 t = twi[((size_t)j << tw_log) + 2 + u];
            s.inv[r2] = CrtInvLds{c.p, t.w, t.ws, c.nr, c.nr_s, c.wr, c.wr_s, 0u};
        }
    }
}

// One wave's group of 64 coefficients starting at g0 (< total = nprod N; a
// multiple of 64): the contraction, the Montgomery finish and the store to
// out.  No workgroup barrier inside.
template <int RT, int KCT, int LT, int INV, int PITCH, int STAGE, class Lds>
__device__ __forceinline__ void crt30_mfma_group(const Lds& s, uint32_t g0, const uint32_t* __restrict__ res,
                                                 uint32_t* __restrict__ out, const int64_t* __restrict__ corr,
                                                 const CrtMfmaConst* __restrict__ mc,
                                                 const QConst* __restrict__ qc, uint32_t log_n, uint32_t L,
                                                 uint32_t k, uint32_t kc_n, uint32_t total, uint32_t q_inv) {
    static_assert(INV == 0 || INV == 2, "fused inverse stages: 0 or 2");
    static_assert(STAGE == 64 || STAGE == 32, "staging: whole wave or one lane half");
    constexpr int LF = 8 * RT - 4;                         // fraction pseudo-limbs LF, LF + 1
    constexpr int LQ = LF < kMaxLimbs ? LF : kMaxLimbs;  // q limbs (zero beyond L)
    constexpr int KCM = KCT > 0 ? KCT : kCrtMaxKC;
    const int8_t* lds_tab = s.tab;
    const CrtInvLds* lds_inv = s.inv;
    if (KCT > 0) kc_n = KCT;
    if (LT > 0) L = LT;
    const uint32_t kbytes = 32 * kc_n;
    // kc_n chunks of 8 primes: k > 8 (KCT - 1), so only the last chunk has rows >= k
    if (KCT > 0 && k <= 8u * (KCT - 1)) __builtin_unreachable();
    const uint32_t N = 1u << log_n;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t col = lane & 31u, half = lane >> 5;
    const int64_t sb_off = 128ll * kbytes;
    // B operands of both tiles of a wave's 64 coefficients starting at g0,
    // every load issued before the first use; rows j >= k read row k - 1 and
    // are replaced by 0 at use
    nc_v4i bf[2][KCM];
    // INV: the wave's first position and the position of this lane's column of a tile
    auto inv_x0 = [&](uint32_t g0) { return ((g0 & (N - 1)) >> 6) * 16u; };
    auto inv_pos = [&](uint32_t x0, int tile, uint32_t c) {
        return x0 + (c & 15u) + (c >> 4) * (N >> 1) + (uint32_t)tile * (N >> 2);
    };
    auto load_group = [&](uint32_t g0) {
#pragma unroll
        for (int tile = 0; tile < 2; ++tile) {
            const uint32_t t0 = g0 + 32u * tile;
            const uint32_t tv = (INV || t0 < total) ? t0 : g0;  // a missing second tile re-reads the first
            const uint32_t prod = uniform((INV ? g0 : tv) >> log_n);
            const char* base = reinterpret_cast<const char*>(res + (size_t)prod * k * N);
            const uint32_t off0 = INV ? inv_pos(inv_x0(g0), tile, col) : (tv & (N - 1)) + col;
            // lane half h reads row 8 kc + 4 h + e: the row 8 kc + e is
            // wave-uniform (a scalar base), the lane's byte offset in it is
            // 4 off0, or 4 (off0 + 4 N) for the upper half.  Where row
            // 8 kc + 4 h + e >= k the lane reads row min(8 kc + e, k - 1)
            // instead (in bounds; such values are replaced by 0 at use).
            const uint32_t lo_b = 4u * off0, hi_b = 4u * (off0 + 4u * N);
            // the row pitch in a register the compiler cannot see through:
            // the 64 row offsets are then recomputed (scalar) at each use
            // rather than hoisted out of the group loop, which they would
            // overflow the scalar registers in
            uint32_t pitch = 4u * N;
            asm volatile("" : "+s"(pitch));
#pragma unroll
            for (int kc = 0; kc < KCM; ++kc) {
                if ((uint32_t)kc < kc_n) {  // wave-uniform
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const uint32_t ra = 8 * kc + e;
                        const uint32_t rowu = ra < k ? ra : k - 1;
                        const uint32_t lb = (half && ra + 4 < k) ? hi_b : lo_b;
                        bf[tile][kc][e] = *reinterpret_cast<const int*>(base + rowu * pitch + lb);
                    }
                }
            }
        }
    };
    load_group(g0);
    if constexpr (INV > 0) {
        // last two inverse stages, residues left in [0, 2p) (rows >= k are
        // replaced by 0 at use; their constants are row k - 1's).  The
        // memory clobber keeps the per-prime constants from being hoisted
        // out of the group loop (32 primes x 9 words would not fit).
        asm volatile("" ::: "memory");
        const uint32_t upper = (col >> 4) & 1u;
#pragma unroll
        for (int kc = 0; kc < KCM; ++kc) {
            if ((uint32_t)kc >= kc_n) continue;  // wave-uniform
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const CrtInvLds ci = lds_inv[2 * (8 * kc + 4 * half + e) + upper];
                const uint32_t p = ci.p, p2 = 2 * p;
                // span N/4: tile 0 (x) with tile 1 (x + N/4)
                uint32_t a = (uint32_t)bf[0][kc][e], b = (uint32_t)bf[1][kc][e];
                Bfly<W30>::gs(a, b, Tw32{ci.w, ci.ws}, p, p2);
                // span N/2 (stage 0) pairs lane c with lane c ^ 16 (rows of
                // 16 lanes: row 2i holds the lower positions, row 2i + 1 the
                // upper).  v_permlane16_swap(a, b) exchanges a's odd rows
                // with b's even rows, so every lane then holds the lower (lo)
                // and upper (hi) value of one tile's pair -- tile 0 in even
                // rows, tile 1 in odd rows -- and forms both outputs; the
                // second swap sends the sums to the even rows and the
                // differences to the odd rows, for both tiles.
                const auto lh = __builtin_amdgcn_permlane16_swap(a, b, false, false);
                const uint32_t lo = lh[0], hi = lh[1];  // [0, 2p) each
                const uint32_t sum = Bfly<W30>::shoup(lo + hi, ci.nr, ci.nr_s, p);
                const uint32_t dif = Bfly<W30>::shoup(lo - hi + p2, ci.wr, ci.wr_s, p);
                const auto sd = __builtin_amdgcn_permlane16_swap(sum, dif, false, false);
                bf[0][kc][e] = (int)sd[0];
                bf[1][kc][e] = (int)sd[1];
            }
        }
    }
    // One row tile: both tiles' contractions (the A fragment serves both),
    // then the exchange.  Lane (c, 0) keeps tile 0's even pseudo-limbs and
    // receives its odd ones; lane (c, 1) receives tile 1's even ones and
    // keeps the odd.  |D| <= Kdim 2^14 < 2^23: byte pairs fit int32.
    auto row_tile = [&](int rt, int32_t (&elo)[4], int32_t (&ehi)[4], int32_t (&olo)[4], int32_t (&ohi)[4],
                        int32_t* dones) {
        nc_v16i acc0 = {}, acc1 = {};
        const int8_t* arow = lds_tab + (rt * 32 + col) * PITCH + 16 * half;
#pragma unroll
        for (int kc = 0; kc < KCM; ++kc) {
            if ((uint32_t)kc >= kc_n) continue;  // wave-uniform
            nc_v4i b0, b1;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const bool live = 8 * kc + 4 * half + e < k;
                b0[e] = live ? bf[0][kc][e] ^ (int)0x80808080 : (int)0x80808080;
                b1[e] = live ? bf[1][kc][e] ^ (int)0x80808080 : (int)0x80808080;
            }
            const nc_v4i a = *reinterpret_cast<const nc_v4i*>(arow + 32 * kc);
            acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b1, acc1, 0, 0, 0);
        }
        if (dones) {  // ones row of pseudo-limb LF + 2 + h, before the swap: own tile
            dones[0] = acc0[12];
            dones[1] = acc1[12];
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t lo0 = (uint32_t)(acc0[4 * m] + (int32_t)((uint32_t)acc0[4 * m + 1] << 8));
            const uint32_t hi0 = (uint32_t)(acc0[4 * m + 2] + (int32_t)((uint32_t)acc0[4 * m + 3] << 8));
            const uint32_t lo1 = (uint32_t)(acc1[4 * m] + (int32_t)((uint32_t)acc1[4 * m + 1] << 8));
            const uint32_t hi1 = (uint32_t)(acc1[4 * m + 2] + (int32_t)((uint32_t)acc1[4 * m + 3] << 8));
            const auto sl = __builtin_amdgcn_permlane32_swap(lo0, lo1, false, false);
            const auto sh = __builtin_amdgcn_permlane32_swap(hi0, hi1, false, false);
            elo[m] = (int32_t)sl[0];
            olo[m] = (int32_t)sl[1];
            ehi[m] = (int32_t)sh[0];
            ohi[m] = (int32_t)sh[1];
        }
    };
    // the last row tile first: it carries the byte-sum row, so every later
    // tile's limbs can be carried into 32-bit limbs as soon as they exist
    int32_t dones[2];
    int32_t lelo[4], lehi[4], lolo[4], lohi[4];
    row_tile(RT - 1, lelo, lehi, lolo, lohi, dones);
    const int64_t sb = (int64_t)(half ? dones[1] : dones[0]) + sb_off;
    const int64_t sbterm = 128 * sb * 0x01010101ll;
    auto ysum = [&](int lam, int32_t lo, int32_t hi) -> int64_t {
        return (int64_t)lo + (int64_t)((uint64_t)(int64_t)hi << 16) + sbterm + corr[lam];
    };
    constexpr int NA = LQ + 2;
    uint32_t x[NA];
    int64_t carry = 0;
#pragma unroll
    for (int rt = 0; rt < RT - 1; ++rt) {
        int32_t elo[4], ehi[4], olo[4], ohi[4];
        row_tile(rt, elo, ehi, olo, ohi, nullptr);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int lam = 8 * rt + i;
            if (lam >= LQ) break;
            const int m = i / 2;
            const int64_t t = ysum(lam, (i & 1) ? olo[m] : elo[m], (i & 1) ? ohi[m] : ehi[m]) + carry;
            x[lam] = (uint32_t)t;
            carry = t >> 32;
        }
    }
    auto last_sum = [&](int lam) -> int64_t {  // pseudo-limb of the last row tile
        const int i = lam - 8 * (RT - 1), m = i / 2;
        return ysum(lam, (i & 1) ? lolo[m] : lelo[m], (i & 1) ? lohi[m] : lehi[m]);
    };
#pragma unroll
    for (int lam = 8 * (RT - 1); lam < LQ; ++lam) {
        const int64_t t = last_sum(lam) + carry;
        x[lam] = (uint32_t)t;
        carry = t >> 32;
    }
    x[LQ] = (uint32_t)carry;  // the sum is non-negative and below 2^(32 L + 17)
#pragma unroll
    for (int l = LQ + 1; l < NA; ++l) x[l] = 0;
    const uint64_t F = (uint64_t)last_sum(LF) + ((uint64_t)last_sum(LF + 1) << 32);
    const uint32_t K = (uint32_t)(F >> 56) + (uint32_t)((F >> 55) & 1u);
    mac_limbs<NA>(x, mc->neg_m_r, LQ, K, 0);
    redc32<LQ, NA>(x, qc, q_inv);
    const uint32_t* r = x + 1;
    // lane (c, h) holds coefficient g0 + 32 h + c = g0 + lane
    uint32_t* st = const_cast<uint32_t*>(s.out) + wave * (uint32_t)STAGE * L;
    if constexpr (STAGE == 64) {
#pragma unroll
        for (int l = 0; l < LQ; ++l)
            if ((uint32_t)l < L) st[lane * L + l] = r[l];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if constexpr (INV > 0) {
            // four runs of 16 coefficients: lanes 16 r .. 16 r + 15 hold run r
            const uint32_t x0 = inv_x0(g0), run_words = 16u * L;  // a multiple of 4
            uint32_t* dst = out + (size_t)(g0 >> log_n) * N * L;
            for (uint32_t w = lane * 4u; w < 64u * L; w += 256u) {
                const uint32_t r = w / run_words, within = w - r * run_words;
                const uint32_t start = x0 + (r & 1u) * (N >> 1) + (r >> 1) * (N >> 2);
                *reinterpret_cast<nc_v4i*>(dst + (size_t)start * L + within) =
                    *reinterpret_cast<const nc_v4i*>(st + w);
            }
        } else {
            const uint32_t words = (total - g0 < 64u ? total - g0 : 64u) * L;  // a multiple of 4 (32 divides the count)
            uint32_t* dst = out + (size_t)g0 * L;
            for (uint32_t w = lane * 4u; w < words; w += 256u)
                *reinterpret_cast<nc_v4i*>(dst + w) = *reinterpret_cast<const nc_v4i*>(st + w);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        // one lane half at a time: the staging holds 32 coefficients, the
        // words of lanes 32 hh .. 32 hh + 31 (32 L of the wave's 64 L)
        static_assert(INV > 0, "half staging is built for the fused form");
        const uint32_t x0 = inv_x0(g0), run_words = 16u * L;  // a multiple of 4
        uint32_t* dst = out + (size_t)(g0 >> log_n) * N * L;
#pragma unroll
        for (uint32_t hh = 0; hh < 2; ++hh) {
            if (half == hh) {
#pragma unroll
                for (int l = 0; l < LQ; ++l)
                    if ((uint32_t)l < L) st[col * L + l] = r[l];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t w = lane * 4u; w < 32u * L; w += 256u) {
                const uint32_t wg = hh * 32u * L + w;  // word of the wave's 64 L
                const uint32_t r = wg / run_words, within = wg - r * run_words;
                const uint32_t start = x0 + (r & 1u) * (N >> 1) + (r >> 1) * (N >> 2);
                *reinterpret_cast<nc_v4i*>(dst + (size_t)start * L + within) =
                    *reinterpret_cast<const nc_v4i*>(st + w);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

// A CRT launch: tables to LDS once, then 256-coefficient groups (one wave
// each quarter) strided over the grid.
template <int RT, int KCT, int LT, int INV = 0>
__device__ __forceinline__ void crt30_mfma_run(const uint32_t* __restrict__ res, uint32_t* __restrict__ out,
                                               const int8_t* __restrict__ tab, const int64_t* __restrict__ corr,
                                               const CrtMfmaConst* __restrict__ mc, const QConst* __restrict__ qc,
                                               uint32_t log_n, uint32_t L, uint32_t k, uint32_t kc_n, uint32_t nprod,
                                               const PrimeConst32* __restrict__ pcs = nullptr,
                                               const Tw32* __restrict__ twi = nullptr, uint32_t tw_log = 0) {
    __shared__ CrtLdsOf<RT, KCT, LT, INV> s;
    crt30_mfma_tables<RT, KCT, LT, INV, kCrtRowStride>(s, tab, k, kc_n, pcs, twi, tw_log);
    __syncthreads();
    const uint32_t total = nprod << log_n;  // a multiple of 32 (N >= 32)
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t q_inv = mc->q_inv;
    for (uint32_t grp = blockIdx.x; grp * 256u < total; grp += gridDim.x) {
        const uint32_t g0 = grp * 256u + wave * 64u;
        if (g0 >= total) continue;  // wave-uniform; no barrier in this loop
        crt30_mfma_group<RT, KCT, LT, INV, kCrtRowStride, 64>(s, g0, res, out, corr, mc, qc, log_n, L, k, kc_n, total,
                                                               q_inv);
    }
}

// The fixed-shape instantiation fits two waves per SIMD without spilling;
// the run-time-shaped one is left to the compiler's register budget.
template <int RT, int KCT, int LT, int INV = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void crt30_mfma_kernel(
    const uint32_t* __restrict__ res, uint32_t* __restrict__ out, const int8_t* __restrict__ tab,
    const int64_t* __restrict__ corr, const CrtMfmaConst* __restrict__ mc, const QConst* __restrict__ qc,
    uint32_t log_n, uint32_t L, uint32_t k, uint32_t kc_n, uint32_t nprod, const PrimeConst32* __restrict__ pcs,
    const Tw32* __restrict__ twi, uint32_t tw_log) {
    crt30_mfma_run<RT, KCT, LT, INV>(res, out, tab, corr, mc, qc, log_n, L, k, kc_n, nprod, pcs, twi, tw_log);
}

template <int RT>
__global__ __launch_bounds__(256) void crt30_mfma_generic_kernel(
    const uint32_t* __restrict__ res, uint32_t* __restrict__ out, const int8_t* __restrict__ tab,
    const int64_t* __restrict__ corr, const CrtMfmaConst* __restrict__ mc, const QConst* __restrict__ qc,
    uint32_t log_n, uint32_t L, uint32_t k, uint32_t kc_n, uint32_t nprod) {
    crt30_mfma_run<RT, 0, 0>(res, out, tab, corr, mc, qc, log_n, L, k, kc_n, nprod);
}

// ===========================================================================
// Persistent batched form (nc_persist.hpp): one launch per batch, every XCD
// draining its own queue of decomposition, NTT and CRT items, so that the
// three stages of different products share every compute unit.  Shape: 30-bit
// primes, both matrix-core paths, outer split 2 (the decomposition runs the
// first two forward stages, the CRT the last two inverse stages), L = 28 and
// 57..64 primes, blocks of N / 4 = 2^LOGB <= 2^12 values.
//
// Hand-off between items: all items of a product run on one XCD (the hardware
// register XCC_ID names it), whose compute units share one L2.  A producer's
// stores are complete in that L2 once its waves pass s_waitcnt vmcnt(0); the
// workgroup then bumps its product's counter with an atomic add.  A consumer
// polls the counter with atomic read-modify-writes (which meet the adds at the
// same place) before it loads.  A T item's compute unit has not read its
// input lines before in this launch (L1 starts empty at every launch) and
// the D items write them as whole 128-byte lines, so any copy its own stores
// left in L1 is current.  A C item's rows were rewritten in place by T items,
// perhaps on other compute units after this one's D item wrote them, so it
// first invalidates its L1 with an agent-scope acquire (buffer_inv sc1).
// Each wait is bounded in time
// (s_memrealtime, 100 MHz) and by an iteration cap: on expiry the workgroup
// sets the XCD's abort flag and error bits and the plan's error word, and
// every workgroup of that XCD stops, so the launch always ends.
// ===========================================================================
constexpr int kPersistCrtPitch = 8 * 32 + 16;  // CRT table pitch for 8 chunks of 8 primes (16-byte padded)
constexpr uint32_t kPersistWaitCap = 1u << 22;  // polls per wait at most (the time bound comes first)

struct PersistArgs {
    const uint32_t* a;  // [batch][N][L] operands
    const uint32_t* b;
    uint32_t* c;        // [batch][N][L] products
    uint32_t* ra;       // [batch][k][N] residue rows of a (the product's rows overwrite them)
    uint32_t* rb;
    uint32_t* ctl;      // nxcd control blocks of xcd_words words (nc_persist.hpp)
    uint32_t* sticky;   // the plan's error word (bits accumulate until the host reads them)
    DevSet30 w;
    uint32_t N, log_n, L, batch, nxcd, xcd_words;
    uint32_t R, nD, nT, nC;
    uint32_t lag;         // rounds between a product's D, T and C items
    uint32_t d_per_poly;  // D items per operand
    uint64_t tmo;       // bound of one readiness wait, s_memrealtime ticks
    uint32_t order[kPersistMaxR];  // one round's slots: kind << 12 | index
};

__device__ __forceinline__ uint32_t xcc_id() {
    // HW_REG_XCC_ID (hardware register 20), bits [3:0]
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);
}

// An atomic read that is a read-modify-write (an add of a zero the compiler
// cannot fold), so it is performed where the producers' atomic adds are.
__device__ __forceinline__ uint32_t atomic_peek(uint32_t* p) {
    uint32_t z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return __hip_atomic_fetch_add(p, z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits (one thread) until *cnt >= need; false on the XCD's abort flag, the
// time bound or the poll cap.
__device__ __forceinline__ bool persist_wait(uint32_t* cnt, uint32_t need, uint32_t* abort_flag, uint64_t tmo) {
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    for (uint32_t poll = 0; poll < kPersistWaitCap; ++poll) {
        if (atomic_peek(cnt) >= need) return true;
        if (atomic_peek(abort_flag)) return false;
        if (__builtin_amdgcn_s_memrealtime() - t0 > tmo) return false;
        __builtin_amdgcn_s_sleep(2);
    }
    return false;
}

template <int LOGB>
// Two waves per SIMD (two workgroups per compute unit, a 256-VGPR budget).
// Built for three, the kernel had 168 VGPRs and spilled about 104 VGPRs and
// 64 SGPRs (400-424 bytes of scratch per lane) across all three item paths,
// and 2-7 % of its launches returned wrong products: one wave's share of a
// D, T or C item's output wrong at a time (decomposition items included,
// which read nothing but the operands and their LDS tables), with no
// readiness-wait error, at three and at two workgroups per compute unit, with
// per-wave agent-scope acquires, and with agent-scope releases too.  At this
// budget it spills 16 VGPRs once at entry (68 bytes) and returned no wrong
// product in 1200 launches with the same schedule (N = 2^14 at batch 128 and
// 256, N = 2^12 at batch 64, two operand sets alternating).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void persist30_kernel(const PersistArgs A) {
    using S = InnerShape<LOGB>;
    static_assert(S::WG == 256 && LOGB <= 12, "one 256-thread workgroup per item, blocks of at most 2^12");
    using CL = CrtLds<4, kPersistCrtPitch, 28, 32, 2 * 8 * 8>;
    union Lds {
        DecompLds<4> d;
        uint2 t[S::BPW * S::PADB];
        CL c;
    };
    __shared__ Lds u;
    __shared__ uint32_t s_next[2];
    __shared__ uint32_t s_fail;
    const uint32_t x = uniform(xcc_id() % A.nxcd);
    uint32_t* xb = A.ctl + (size_t)x * A.xcd_words;
    const uint32_t m_max = (A.batch + A.nxcd - 1) / A.nxcd;
    uint32_t* dcnt = xb + kCtlCounts;
    uint32_t* tcnt = xb + kCtlCounts + m_max;
    const uint32_t m = persist_products_on(A.batch, A.nxcd, x);
    const uint32_t slots = persist_slots(A.R, m, A.lag);
    const uint32_t N = A.N, k = A.w.k, L = A.L;
    const uint32_t total = A.batch << A.log_n;
    const uint32_t q_inv = A.w.crt_mc->q_inv;
    const uint32_t sub = threadIdx.x / S::TPB, tid = threadIdx.x % S::TPB;
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t role = 3;  // the kind whose tables LDS holds (3: none)
    uint32_t done = 0;
    if (threadIdx.x == 0) {
        s_next[0] = atomicAdd(&xb[kCtlHead], 1u);
        s_fail = 0;
    }
    for (uint32_t iter = 0;; ++iter) {
        __syncthreads();  // the previous item's LDS and s_next reads are done
        const uint32_t it = uniform(s_next[iter & 1]);
        if (it >= slots || s_fail) break;
        uint32_t nxt = 0;  // the next queue position, fetched while this item runs
        if (threadIdx.x == 0) nxt = atomicAdd(&xb[kCtlHead], 1u);
        const PersistItem item = persist_decode(it, A.R, A.order, m, A.lag);
        if (item.valid) {
            const uint32_t kind = uniform(item.kind), idx = uniform(item.idx), lp = uniform(item.local);
            const uint32_t P = x + A.nxcd * lp;  // the product
            if (kind != kItemD) {
                if (threadIdx.x == 0) {
                    const bool ok = kind == kItemT ? persist_wait(&dcnt[lp], A.nD, &xb[kCtlAbort], A.tmo)
                                                   : persist_wait(&tcnt[lp], A.nT, &xb[kCtlAbort], A.tmo);
                    if (!ok) {
                        atomicOr(&xb[kCtlErr], kPersistErrTimeout);
                        atomicOr(A.sticky, kPersistErrTimeout);
                        __hip_atomic_fetch_or(&xb[kCtlAbort], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        s_fail = 1;
                    } else if (kind == kItemC) {
                        // a C item's rows were rewritten in place by T items
                        // on other compute units: drop this one's L1 copies
                        // (an agent-scope acquire, buffer_inv sc1; the
                        // workgroup-scope form does not reach other units'
                        // writes)
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    }
                }
                __syncthreads();
                if (s_fail) break;
            }
            if (kind == kItemD) {
                if (role != kItemD) {
                    decompose30_mfma_tables<2>(u.d, A.w.mfma_tab, A.w.mfma_bias, A.w.pc, A.w.tw_fwd, k,
                                               A.w.mfma_tiles, A.w.tw_log);
                    __syncthreads();
                    role = kItemD;
                }
                const uint32_t poly = idx / A.d_per_poly;  // 0: a, 1: b
                const uint32_t chunk0 = (idx - poly * A.d_per_poly) * 4u;
                const uint32_t* src = (poly ? A.b : A.a) + (size_t)P * N * L;
                uint32_t* dst = (poly ? A.rb : A.ra) + (size_t)P * k * N;
                decompose30_mfma_item<2, 1>(u.d, src, dst, chunk0, N, L, k, A.w.mfma_tiles);
            } else if (kind == kItemT) {
                role = kItemT;  // the exchanges overwrite any tables
                uint32_t wid = idx * S::BPW + sub;  // (prime, block) of the product, 4 blocks per prime
                if constexpr (S::TPB >= 64) wid = uniform(wid);
                const bool active = wid < 4u * k;
                if (!active) wid = 0;
                const uint32_t j = wid >> 2, blk = wid & 3u;
                const size_t off = ((size_t)P * k + j) * N + ((size_t)blk << LOGB);
                const PrimeConst32 c = A.w.pc[merge30(LOGB) * k + j];  // the block's merge variant
                ntt_inner30_block<LOGB>(u.t + sub * S::PADB, tid, A.ra + off, A.rb + off, blk << LOGB, c,
                                        A.w.tw_fwd + ((size_t)j << A.w.tw_log),
                                        A.w.tw_inv + ((size_t)j << A.w.tw_log), A.log_n, active, false);
            } else {
                if (role != kItemC) {
                    crt30_mfma_tables<4, 8, 28, 2, kPersistCrtPitch>(u.c, A.w.crt_tab, k, 8u,
                                                                    A.w.pc + merge30(LOGB) * k, A.w.tw_inv,
                                                                    A.w.tw_log);
                    __syncthreads();
                    role = kItemC;
                }
                const uint32_t pos = idx * 256u + wave * 64u;  // the wave's first position in the product
                if (pos < N)
                    crt30_mfma_group<4, 8, 28, 2, kPersistCrtPitch, 32>(u.c, P * N + pos, A.ra, A.c, A.w.crt_c,
                                                                       A.w.crt_mc, A.w.qc, A.log_n, L, k, 8u, total,
                                                                       q_inv);
            }
            if (kind != kItemC) {  // publish: this item's stores are in L2, then count it
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __syncthreads();
                if (threadIdx.x == 0)
                    __hip_atomic_fetch_add(kind == kItemD ? &dcnt[lp] : &tcnt[lp], 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
            ++done;
        }
        if (threadIdx.x == 0) s_next[(iter + 1) & 1] = nxt;
    }
    if (threadIdx.x == 0) atomicAdd(&xb[kCtlDone], done);
}

__global__ __launch_bounds__(256) void persist_reset_kernel(uint32_t* __restrict__ ctl, uint32_t words) {
    for (uint32_t i = threadIdx.x; i < words; i += 256) ctl[i] = 0u;
}

// ===========================================================================
// Twiddle tables on the device (what fherma_init builds on the host,
// solve.cu:1224-1257, here in the merged-psi form): entry x of prime j is
// psi_j^(+-bitrev(x)) with its Shoup quotient, {0, 0} at x = 0.  One thread
// per (prime, x): the power is a chain of Shoup products over the set bits
// of bitrev(x) with the table pow[j][dir][i] = psi^(+-2^i) (lazy in [0, 2p),
// reduced once), and the quotient floor(w 2^b / p) is exact long division.
// ===========================================================================
__device__ __forceinline__ uint64_t shoup_quotient64(uint64_t w, uint64_t p) {
    uint64_t r = w, q = 0;  // w < p < 2^62: 2 r stays below 2^63
#pragma unroll 8
    for (int i = 0; i < 64; ++i) {
        r <<= 1;
        q <<= 1;
        if (r >= p) {
            r -= p;
            q |= 1;
        }
    }
    return q;
}

__global__ __launch_bounds__(256) void twiddles62_kernel(Tw* __restrict__ fwd, Tw* __restrict__ inv,
                                                         const Tw* __restrict__ pow, const PrimeConst* __restrict__ pc,
                                                         uint32_t log_n, uint32_t k) {
    const uint32_t N = 1u << log_n;
    const uint64_t gid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (gid >= (uint64_t)k * N) return;
    const uint32_t j = (uint32_t)(gid >> log_n), x = (uint32_t)gid & (N - 1);
    Tw f{0, 0}, iv{0, 0};
    if (x != 0) {
        const uint64_t p = pc[j].p;
        const uint32_t e = __builtin_bitreverse32(x) >> (32 - log_n);
        const Tw* pf = pow + (size_t)j * 2 * log_n;
        const Tw* pi = pf + log_n;
        uint64_t a = 1, b = 1;
        for (uint32_t i = 0; i < log_n; ++i) {
            if ((e >> i) & 1u) {
                a = W62::shoup(a, pf[i].w, pf[i].ws, p);
                b = W62::shoup(b, pi[i].w, pi[i].ws, p);
            }
        }
        a = W62::sub_if_ge(a, p);
        b = W62::sub_if_ge(b, p);
        f = Tw{a, shoup_quotient64(a, p)};
        iv = Tw{b, shoup_quotient64(b, p)};
    }
    fwd[gid] = f;
    inv[gid] = iv;
}

// 30-bit set: the forward value is stored negated (2^32 - w), as the
// forward butterfly expects.
__global__ __launch_bounds__(256) void twiddles30_kernel(Tw32* __restrict__ fwd, Tw32* __restrict__ inv,
                                                         const Tw32* __restrict__ pow,
                                                         const PrimeConst32* __restrict__ pc, uint32_t log_n,
                                                         uint32_t k) {
    const uint32_t N = 1u << log_n;
    const uint64_t gid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (gid >= (uint64_t)k * N) return;
    const uint32_t j = (uint32_t)(gid >> log_n), x = (uint32_t)gid & (N - 1);
    Tw32 f{0u, 0u}, iv{0u, 0u};
    if (x != 0) {
        const uint32_t p = pc[j].p;
        const uint32_t e = __builtin_bitreverse32(x) >> (32 - log_n);
        const Tw32* pf = pow + (size_t)j * 2 * log_n;
        const Tw32* pi = pf + log_n;
        uint32_t a = 1, b = 1;
        for (uint32_t i = 0; i < log_n; ++i) {
            if ((e >> i) & 1u) {
                a = W30::shoup(a, pf[i].w, pf[i].ws, p);
                b = W30::shoup(b, pi[i].w, pi[i].ws, p);
            }
        }
        a = W30::sub_if_ge(a, p);
        b = W30::sub_if_ge(b, p);
        f = Tw32{0u - a, (uint32_t)(((uint64_t)a << 32) / p)};
        iv = Tw32{b, (uint32_t)(((uint64_t)b << 32) / p)};
    }
    fwd[gid] = f;
    inv[gid] = iv;
}

// ===========================================================================
// 32x32->64 multiply throughput probe: eight independent v_mad_u64_u32 chains
// per lane, no other VALU work in the loop.
// ===========================================================================
__global__ __launch_bounds__(256) void mul32_peak_kernel(const uint32_t* __restrict__ in, uint64_t* __restrict__ out,
                                                         uint32_t iters) {
    const uint32_t b = in[threadIdx.x & 63u] | 1u;
    const uint64_t c = in[(threadIdx.x + 7u) & 63u];
    uint64_t acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = in[(threadIdx.x + (uint32_t)u) & 63u];
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] = (uint64_t)(uint32_t)acc[u] * b + c;
    }
    uint64_t x = 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) x ^= acc[u];
    out[blockIdx.x * 256 + threadIdx.x] = x;
}

// ===========================================================================
// launchers
// ===========================================================================
uint32_t max_log_blocks(uint32_t log_n) {
    if (log_n < 4) return 0;
    const uint32_t m = log_n - 4;  // a block keeps >= 16 values
    return m < (uint32_t)kMaxOuterBits ? m : (uint32_t)kMaxOuterBits;
}

uint32_t min_log_blocks(uint32_t log_n) { return log_n > (uint32_t)kMaxLogBlock ? log_n - kMaxLogBlock : 0u; }

template <class A>
struct SetOf;
template <>
struct SetOf<W62> {
    static const DevSet62& get(const DevPlan& d) { return d.w62; }
};
template <>
struct SetOf<W30> {
    static const DevSet30& get(const DevPlan& d) { return d.w30; }
};

// The prime constants the final inverse stage reads after an inner block of
// 2^logb: for the 30-bit set the variant of that block's merge depth
// (merge30; 0 for the one-operand and tiny kernels, whose blocks are below
// 2^5), for the 62-bit set the only one.
template <class A>
static const typename A::Prime* pc_for(const DevPlan& d, uint32_t logb) {
    const auto& set = SetOf<A>::get(d);
    if constexpr (std::is_same<A, W30>::value) return set.pc + (size_t)merge30(logb) * set.k;
    return set.pc;
}

template <class A, int LOGB>
static void launch_inner(const DevPlan& d, void* ra, const void* rb, uint32_t nprod, hipStream_t s) {
    using S = InnerShape<LOGB>;
    using T = typename A::T;
    const auto& set = SetOf<A>::get(d);
    const uint32_t items = (set.k * nprod) << (d.log_n - LOGB);
    const uint32_t grid = (items + S::BPW - 1) / S::BPW;
    if constexpr (std::is_same<A, W30>::value && LOGB >= 5)
        ntt_inner30_kernel<LOGB><<<grid, S::WG, 0, s>>>(static_cast<T*>(ra), static_cast<const T*>(rb), set.tw_fwd,
                                                         set.tw_inv, pc_for<A>(d, LOGB), d.log_n, set.k, nprod,
                                                         set.tw_log);
    else
        ntt_inner_kernel<A, LOGB><<<grid, S::WG, 0, s>>>(static_cast<T*>(ra), static_cast<const T*>(rb), set.tw_fwd,
                                                          set.tw_inv, pc_for<A>(d, LOGB), d.log_n, set.k, nprod,
                                                          set.tw_log);
}

template <class A>
static hipError_t inner_dispatch(const DevPlan& d, void* ra, const void* rb, uint32_t nprod, uint32_t log_blocks,
                                 hipStream_t s) {
    using T = typename A::T;
    const auto& set = SetOf<A>::get(d);
    if (d.log_n < 4) {
        const uint32_t t = set.k * nprod;
        ntt_tiny_kernel<A><<<(t + 63) / 64, 64, 0, s>>>(static_cast<T*>(ra), static_cast<const T*>(rb), set.tw_fwd,
                                                        set.tw_inv, set.pc, d.log_n, set.k, nprod, set.tw_log);
        return hipGetLastError();
    }
    switch (d.log_n - log_blocks) {
        case 4: launch_inner<A, 4>(d, ra, rb, nprod, s); break;
        case 5: launch_inner<A, 5>(d, ra, rb, nprod, s); break;
        case 6: launch_inner<A, 6>(d, ra, rb, nprod, s); break;
        case 7: launch_inner<A, 7>(d, ra, rb, nprod, s); break;
        case 8: launch_inner<A, 8>(d, ra, rb, nprod, s); break;
        case 9: launch_inner<A, 9>(d, ra, rb, nprod, s); break;
        case 10: launch_inner<A, 10>(d, ra, rb, nprod, s); break;
        case 11: launch_inner<A, 11>(d, ra, rb, nprod, s); break;
        case 12: launch_inner<A, 12>(d, ra, rb, nprod, s); break;
        case 13: launch_inner<A, 13>(d, ra, rb, nprod, s); break;
        case 14: launch_inner<A, 14>(d, ra, rb, nprod, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class A, int S1>
static void launch_outer_fwd(const DevPlan& d, void* ra, void* rb, uint32_t nprod, hipStream_t s) {
    using T = typename A::T;
    const auto& set = SetOf<A>::get(d);
    const uint32_t wgs_per_row = ((d.N >> S1) + 255) / 256;
    const uint32_t grid = wgs_per_row * 2 * set.k * nprod;
    ntt_outer_fwd_kernel<A, S1><<<grid, 256, 0, s>>>(static_cast<T*>(ra), static_cast<T*>(rb), set.tw_fwd, set.pc,
                                                      d.log_n, set.k, nprod, set.tw_log);
}

template <class A, int S1>
static void launch_outer_inv(const DevPlan& d, void* ra, uint32_t nprod, hipStream_t s) {
    using T = typename A::T;
    const auto& set = SetOf<A>::get(d);
    const uint32_t wgs_per_row = ((d.N >> S1) + 255) / 256;
    const uint32_t grid = wgs_per_row * set.k * nprod;
    ntt_outer_inv_kernel<A, S1><<<grid, 256, 0, s>>>(static_cast<T*>(ra), set.tw_inv, pc_for<A>(d, d.log_n - S1),
                                                      d.log_n, set.k, nprod, set.tw_log);
}

template <class A>
static hipError_t outer_fwd_dispatch(const DevPlan& d, void* ra, void* rb, uint32_t nprod, uint32_t log_blocks,
                                     hipStream_t s) {
    if (d.log_n < 4) return hipSuccess;
    switch (log_blocks) {
        case 0: break;
        case 1: launch_outer_fwd<A, 1>(d, ra, rb, nprod, s); break;
        case 2: launch_outer_fwd<A, 2>(d, ra, rb, nprod, s); break;
        case 3: launch_outer_fwd<A, 3>(d, ra, rb, nprod, s); break;
        case 4: launch_outer_fwd<A, 4>(d, ra, rb, nprod, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class A>
static hipError_t outer_inv_dispatch(const DevPlan& d, void* ra, uint32_t nprod, uint32_t log_blocks,
                                     hipStream_t s) {
    if (d.log_n < 4) return hipSuccess;
    switch (log_blocks) {
        case 0: break;
        case 1: launch_outer_inv<A, 1>(d, ra, nprod, s); break;
        case 2: launch_outer_inv<A, 2>(d, ra, nprod, s); break;
        case 3: launch_outer_inv<A, 3>(d, ra, nprod, s); break;
        case 4: launch_outer_inv<A, 4>(d, ra, nprod, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool decompose_fuses_outer(const DevPlan& d, int width, bool mfma, uint32_t log_blocks) {
    return width == 30 && mfma && d.w30.mfma_tab && log_blocks >= 1 && log_blocks <= 2 && d.log_n >= 4 &&
           (d.N >> log_blocks) >= 32;
}

hipError_t launch_decompose(const DevPlan& d, int width, const uint32_t* a, void* res_a, uint32_t na,
                            const uint32_t* b, void* res_b, uint32_t nb, hipStream_t s, bool mfma,
                            uint32_t fuse_s1) {
    if (fuse_s1 && !decompose_fuses_outer(d, width, mfma, fuse_s1)) return hipErrorInvalidValue;
    const dim3 grid(((d.N + 63) / 64) * (na + nb));
    if (width == 62) {
        uint64_t* ra = static_cast<uint64_t*>(res_a);
        uint64_t* rb = static_cast<uint64_t*>(res_b);
        if (d.L == 28)
            decompose62_kernel<28><<<grid, 256, 0, s>>>(a, ra, na, b, rb, d.w62.place, d.w62.pc, d.N, d.L, d.w62.k);
        else
            decompose62_kernel<kMaxLimbs><<<grid, 256, 0, s>>>(a, ra, na, b, rb, d.w62.place, d.w62.pc, d.N, d.L,
                                                               d.w62.k);
    } else if (mfma && d.w30.mfma_tab && d.N >= 32) {
        constexpr uint32_t kPerWg = 4 * kMfmaCoeffTiles * 32;
        const uint32_t items1 = ((d.N + kPerWg - 1) / kPerWg) * (na + nb);
        // fewer work items than CUs (or than two per CU): 4 (2) waves per chunk
        const uint32_t ws = items1 < d.cus ? 4u : items1 < 2 * d.cus ? 2u : 1u;
        const uint32_t items = items1 * ws;
        const dim3 g(items < 3 * d.cus ? items : 3 * d.cus);  // persistent: three workgroups per CU
        uint32_t* ra = static_cast<uint32_t*>(res_a);
        uint32_t* rb = static_cast<uint32_t*>(res_b);
        const DevSet30& w = d.w30;
        auto go = [&](auto kern) {
            kern<<<g, 256, 0, s>>>(a, ra, na, b, rb, w.mfma_tab, w.mfma_bias, w.pc, w.tw_fwd, d.N, d.L, w.k,
                                   w.mfma_tiles, nb, w.tw_log);
        };
        switch (fuse_s1 * 3 + (ws == 1 ? 0 : ws == 2 ? 1 : 2)) {
            case 0: go(decompose30_mfma_kernel<0, 1>); break;
            case 1: go(decompose30_mfma_kernel<0, 2>); break;
            case 2: go(decompose30_mfma_kernel<0, 4>); break;
            case 3: go(decompose30_mfma_kernel<1, 1>); break;
            case 4: go(decompose30_mfma_kernel<1, 2>); break;
            case 5: go(decompose30_mfma_kernel<1, 4>); break;
            case 6: go(decompose30_mfma_kernel<2, 1>); break;
            case 7: go(decompose30_mfma_kernel<2, 2>); break;
            case 8: go(decompose30_mfma_kernel<2, 4>); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        uint32_t* ra = static_cast<uint32_t*>(res_a);
        uint32_t* rb = static_cast<uint32_t*>(res_b);
        if (d.L == 28)
            decompose30_kernel<28><<<grid, 256, 0, s>>>(a, ra, na, b, rb, d.w30.place, d.w30.pc, d.N, d.L, d.w30.k);
        else
            decompose30_kernel<kMaxLimbs><<<grid, 256, 0, s>>>(a, ra, na, b, rb, d.w30.place, d.w30.pc, d.N, d.L,
                                                               d.w30.k);
    }
    return hipGetLastError();
}

hipError_t launch_ntt_outer_fwd(const DevPlan& d, int width, void* ra, void* rb, uint32_t nprod, uint32_t log_blocks,
                                hipStream_t s) {
    return width == 62 ? outer_fwd_dispatch<W62>(d, ra, rb, nprod, log_blocks, s)
                       : outer_fwd_dispatch<W30>(d, ra, rb, nprod, log_blocks, s);
}

hipError_t launch_ntt_inner(const DevPlan& d, int width, void* ra, const void* rb, uint32_t nprod,
                            uint32_t log_blocks, hipStream_t s) {
    return width == 62 ? inner_dispatch<W62>(d, ra, rb, nprod, log_blocks, s)
                       : inner_dispatch<W30>(d, ra, rb, nprod, log_blocks, s);
}

hipError_t launch_ntt_outer_inv(const DevPlan& d, int width, void* ra, uint32_t nprod, uint32_t log_blocks,
                                hipStream_t s) {
    return width == 62 ? outer_inv_dispatch<W62>(d, ra, nprod, log_blocks, s)
                       : outer_inv_dispatch<W30>(d, ra, nprod, log_blocks, s);
}

hipError_t launch_ntt_product(const DevPlan& d, int width, void* ra, void* rb, uint32_t nprod, uint32_t log_blocks,
                              hipStream_t s, bool outer_fwd_done, bool outer_inv_in_crt) {
    if (d.log_n >= 4 && (log_blocks < min_log_blocks(d.log_n) || log_blocks > max_log_blocks(d.log_n)))
        return hipErrorInvalidValue;
    hipError_t e = outer_fwd_done ? hipSuccess : launch_ntt_outer_fwd(d, width, ra, rb, nprod, log_blocks, s);
    if (e != hipSuccess) return e;
    e = launch_ntt_inner(d, width, ra, rb, nprod, log_blocks, s);
    if (e != hipSuccess || outer_inv_in_crt) return e;
    return launch_ntt_outer_inv(d, width, ra, nprod, log_blocks, s);
}

bool crt_mfma_available(const DevPlan& d) { return d.w30.crt_tab != nullptr && d.N >= 32; }

template <int RT, int KCT = 0, int LT = 0>
static void launch_crt_mfma(const DevPlan& d, const uint32_t* res, uint32_t* out, uint32_t nprod, hipStream_t s) {
    const uint32_t groups = ((nprod << d.log_n) + 255) / 256;
    const uint32_t lds = RT * 32u * kCrtRowStride + 4u * 64u * 4u * (LT > 0 ? LT : kMaxLimbs);
    const uint32_t per_cu = (64u * 1024u * 5u / 2u) / lds;  // workgroups whose LDS fits 160 KiB
    const uint32_t cap = d.cus * (per_cu < 2u ? per_cu : 2u);  // two waves per SIMD at most
    const uint32_t grid = groups < cap ? groups : cap;
    if constexpr (KCT > 0)
        crt30_mfma_kernel<RT, KCT, LT><<<grid, 256, 0, s>>>(res, out, d.w30.crt_tab, d.w30.crt_c, d.w30.crt_mc, d.w30.qc,
                                                         d.log_n, d.L, d.w30.k, d.w30.crt_kc, nprod, nullptr,
                                                         nullptr, 0u);
    else
        crt30_mfma_generic_kernel<RT><<<grid, 256, 0, s>>>(res, out, d.w30.crt_tab, d.w30.crt_c, d.w30.crt_mc,
                                                            d.w30.qc, d.log_n, d.L, d.w30.k, d.w30.crt_kc, nprod);
}

bool crt_fuses_outer_inv(const DevPlan& d, int width, bool mfma, uint32_t log_blocks) {
    return width == 30 && mfma && crt_mfma_available(d) && log_blocks == 2 && d.log_n >= 6 && d.w30.crt_rt == 4 &&
           d.w30.crt_kc == 8 && d.L == 28;
}

hipError_t launch_crt_inv2(const DevPlan& d, const uint32_t* res, uint32_t* out, uint32_t nprod, hipStream_t s) {
    if (!crt_fuses_outer_inv(d, 30, true, 2)) return hipErrorInvalidValue;
    constexpr uint32_t lds = 4u * 32u * kCrtRowStride + 4u * 64u * 4u * 28u;
    const uint32_t groups = ((nprod << d.log_n) + 255) / 256;
    // + the per-lane-half inverse-stage constants of 8 chunks of 8 primes
    const uint32_t per_cu = (64u * 1024u * 5u / 2u) / (lds + (uint32_t)(2 * 8 * 8 * sizeof(CrtInvLds)));
    const uint32_t cap = d.cus * (per_cu < 2u ? per_cu : 2u);
    const uint32_t grid = groups < cap ? groups : cap;
    crt30_mfma_kernel<4, 8, 28, 2><<<grid, 256, 0, s>>>(res, out, d.w30.crt_tab, d.w30.crt_c, d.w30.crt_mc, d.w30.qc,
                                                        d.log_n, d.L, d.w30.k, d.w30.crt_kc, nprod,
                                                        pc_for<W30>(d, d.log_n - 2), d.w30.tw_inv, d.w30.tw_log);
    return hipGetLastError();
}

// The persistent form is launched (set false to route every call through
// the launch sequence).
constexpr bool kPersistLaunchEnabled = true;

bool persist_launch_enabled() { return kPersistLaunchEnabled; }

bool persist_available(const DevPlan& d, uint32_t batch) {
    return kPersistLaunchEnabled && d.xcds >= 1 && d.xcds <= kPersistMaxXcd && batch >= d.xcds && d.log_n >= 10 && d.log_n <= 14 &&
           crt_fuses_outer_inv(d, 30, true, 2) && decompose_fuses_outer(d, 30, true, 2) && d.w30.k <= 64 &&
           persist_shape_for(d.log_n, d.w30.k).R <= kPersistMaxR;
}

size_t persist_ctl_bytes(const DevPlan& d, uint32_t batch) {
    const uint32_t nxcd = d.xcds ? d.xcds : 1u;
    return (size_t)nxcd * persist_xcd_words(batch, nxcd) * sizeof(uint32_t);
}

hipError_t launch_persist30(const DevPlan& d, const uint32_t* a, const uint32_t* b, uint32_t* c, uint32_t* ra,
                            uint32_t* rb, uint32_t batch, uint32_t* ctl, uint32_t* sticky, uint64_t tmo,
                            hipStream_t s) {
    if (!kPersistLaunchEnabled) return hipErrorNotSupported;
    if (!persist_available(d, batch)) return hipErrorInvalidValue;
    const PersistShape sh = persist_shape_for(d.log_n, d.w30.k);
    PersistArgs A{};
    A.a = a;
    A.b = b;
    A.c = c;
    A.ra = ra;
    A.rb = rb;
    A.ctl = ctl;
    A.sticky = sticky;
    A.w = d.w30;
    A.N = d.N;
    A.log_n = d.log_n;
    A.L = d.L;
    A.batch = batch;
    A.nxcd = d.xcds;
    A.xcd_words = persist_xcd_words(batch, d.xcds);
    A.R = sh.R;
    A.nD = sh.nD;
    A.nT = sh.nT;
    A.nC = sh.nC;
    A.d_per_poly = sh.nD / 2;
    A.tmo = tmo;
    persist_order(sh, A.order);
    const uint32_t words = d.xcds * A.xcd_words;
    persist_reset_kernel<<<1, 256, 0, s>>>(ctl, words);
    // two workgroups per compute unit (53 KB of LDS and up to 256 VGPRs
    // each); workgroups go round-robin over the XCDs, so each XCD's queue gets
    // 2 cus / xcds of them
    const dim3 grid(2 * d.cus);
    A.lag = persist_lag(sh.R, grid.x / d.xcds);
    switch (d.log_n) {
        case 10: persist30_kernel<8><<<grid, 256, 0, s>>>(A); break;
        case 11: persist30_kernel<9><<<grid, 256, 0, s>>>(A); break;
        case 12: persist30_kernel<10><<<grid, 256, 0, s>>>(A); break;
        case 13: persist30_kernel<11><<<grid, 256, 0, s>>>(A); break;
        case 14: persist30_kernel<12><<<grid, 256, 0, s>>>(A); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_crt(const DevPlan& d, int width, const void* res, uint32_t* out, uint32_t nprod, hipStream_t s,
                      bool mfma) {
    const uint32_t total = nprod << d.log_n;
    const dim3 grid((total + 255) / 256);
    if (width == 30 && mfma && crt_mfma_available(d)) {
        const uint32_t* r = static_cast<const uint32_t*>(res);
        switch (d.w30.crt_rt) {
            case 1: launch_crt_mfma<1>(d, r, out, nprod, s); break;
            case 2: launch_crt_mfma<2>(d, r, out, nprod, s); break;
            case 3: launch_crt_mfma<3>(d, r, out, nprod, s); break;
            case 4:
                if (d.w30.crt_kc == 8 && d.L == 28) launch_crt_mfma<4, 8, 28>(d, r, out, nprod, s);  // 57..64 primes
                else launch_crt_mfma<4>(d, r, out, nprod, s);
                break;
            case 5: launch_crt_mfma<5>(d, r, out, nprod, s); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (width == 62) {
        const uint64_t* r = static_cast<const uint64_t*>(res);
        if (d.L == 28)
            crt62_kernel<28><<<grid, 256, 0, s>>>(r, out, d.w62.cc, d.w62.miq, d.w62.qc, d.w62.pc, d.log_n, d.L,
                                                  d.w62.k, nprod);
        else
            crt62_kernel<kMaxLimbs><<<grid, 256, 0, s>>>(r, out, d.w62.cc, d.w62.miq, d.w62.qc, d.w62.pc, d.log_n,
                                                         d.L, d.w62.k, nprod);
    } else {
        const uint32_t* r = static_cast<const uint32_t*>(res);
        if (d.L == 28)
            crt30_kernel<28><<<grid, 256, 0, s>>>(r, out, d.w30.cc, d.w30.miq, d.w30.qc, d.w30.pc, d.log_n, d.L,
                                                  d.w30.k, nprod);
        else
            crt30_kernel<kMaxLimbs><<<grid, 256, 0, s>>>(r, out, d.w30.cc, d.w30.miq, d.w30.qc, d.w30.pc, d.log_n,
                                                         d.L, d.w30.k, nprod);
    }
    return hipGetLastError();
}

// ===========================================================================
// Table-set initialisation on the device (the work fherma_init does on the
// host, solve.cu:1174-1204 and 1342-1369): the prime search with each
// prime's psi, and the big-integer CRT tables mod q.
// ===========================================================================

// Montgomery arithmetic mod an odd n < 2^63 (R = 2^64), for the prime search.
struct Mont64 {
    uint64_t n, ninv;  // ninv = -n^-1 mod 2^64
    __device__ explicit Mont64(uint64_t n_) : n(n_) {
        uint64_t x = n_;  // Newton: x = n^-1 mod 2^(2^i), 6 steps from 3 bits
        for (int i = 0; i < 6; ++i) x *= 2 - n_ * x;
        ninv = 0 - x;
    }
    __device__ uint64_t mul(uint64_t a, uint64_t b) const {  // a b R^-1 mod n, a, b < n
        const uint64_t lo = a * b, hi = __umul64hi(a, b);
        const uint64_t m = lo * ninv;
        const uint64_t mlo = m * n, mhi = __umul64hi(m, n);
        const uint64_t c = (lo + mlo) < lo ? 1u : 0u;  // carry of the low words (their sum is 0 mod 2^64)
        const uint64_t r = hi + mhi + c;               // < 2n
        return r >= n ? r - n : r;
    }
    __device__ uint64_t to(uint64_t a) const {  // a R mod n
        return (uint64_t)(((unsigned __int128)(a % n) << 64) % n);
    }
    __device__ uint64_t pow(uint64_t a_m, uint64_t e) const {  // a_m in Montgomery form; result too
        uint64_t r = to(1), b = a_m;
        for (; e; e >>= 1) {
            if (e & 1) r = mul(r, b);
            b = mul(b, b);
        }
        return r;
    }
};

// Deterministic Miller-Rabin for n < 3.3e24 (the first twelve prime bases).
__device__ bool mr_prime(uint64_t n) {
    if (n < 2) return false;
    const uint32_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint32_t b : bases)
        if (n % b == 0) return n == b;
    const Mont64 M(n);
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++s;
    }
    const uint64_t one = M.to(1), minus_one = M.to(n - 1);
    for (uint32_t b : bases) {
        uint64_t x = M.pow(M.to(b), d);
        if (x == one || x == minus_one) continue;
        bool witness = true;
        for (int r = 1; r < s && witness; ++r) {
            x = M.mul(x, x);
            if (x == minus_one) witness = false;
        }
        if (witness) return false;
    }
    return true;
}

// Candidate i is top - i step (top = 1 mod step).  out[i] = 0 if it is not
// prime, else psi = g^((p - 1) / step) for the smallest g >= 2 with
// g^((p - 1) / 2) = -1 (a primitive step-th root of unity: the same psi rule
// as the host's, solve.cu:495-502).
__global__ __launch_bounds__(64) void prime_search_kernel(uint64_t top, uint64_t step, uint32_t count,
                                                          uint64_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= count) return;
    const uint64_t p = top - (uint64_t)i * step;
    uint64_t psi = 0;
    if (p > step && mr_prime(p)) {
        const Mont64 M(p);
        const uint64_t minus_one = M.to(p - 1);
        for (uint64_t g = 2; g < 100000; ++g) {
            const uint64_t gm = M.to(g);
            if (M.pow(gm, (p - 1) / 2) == minus_one) {
                psi = M.mul(M.pow(gm, (p - 1) / step), 1);  // out of Montgomery form
                break;
            }
        }
    }
    out[i] = psi;
}

// ---- L-limb arithmetic mod q (x < q throughout), one big number per thread
__device__ bool limbs_geq(const uint32_t* x, const uint32_t* q, uint32_t L) {
    for (int l = (int)L - 1; l >= 0; --l)
        if (x[l] != q[l]) return x[l] > q[l];
    return true;
}

__device__ void limbs_sub(uint32_t* x, const uint32_t* q, uint32_t L) {
    uint32_t br = 0;
    for (uint32_t l = 0; l < L; ++l) {
        uint32_t nb;
        x[l] = __builtin_subc(x[l], q[l], br, &nb);
        br = nb;
    }
}

__device__ void dbl_mod(uint32_t* x, const uint32_t* q, uint32_t L) {  // x = 2 x mod q
    uint32_t top = 0;
    for (uint32_t l = 0; l < L; ++l) {
        const uint32_t nt = x[l] >> 31;
        x[l] = (x[l] << 1) | top;
        top = nt;
    }
    if (top || limbs_geq(x, q, L)) limbs_sub(x, q, L);  // 2x < 2q
}

__device__ void add_mod(uint32_t* x, const uint32_t* y, const uint32_t* q, uint32_t L) {  // x = x + y mod q
    uint32_t c = 0;
    for (uint32_t l = 0; l < L; ++l) {
        uint32_t nc;
        x[l] = __builtin_addc(x[l], y[l], c, &nc);
        c = nc;
    }
    if (c || limbs_geq(x, q, L)) limbs_sub(x, q, L);
}

__device__ void mul_word_mod(uint32_t* x, uint64_t w, const uint32_t* q, uint32_t L) {  // x = x w mod q
    uint32_t acc[kMaxLimbs];
    for (uint32_t l = 0; l < L; ++l) acc[l] = 0;
    for (int bit = 63; bit >= 0; --bit) {
        dbl_mod(acc, q, L);
        if ((w >> bit) & 1u) add_mod(acc, x, q, L);
    }
    for (uint32_t l = 0; l < L; ++l) x[l] = acc[l];
}

// Thread j < k: (M / p_j) mod q = prod_{i != j} p_i mod q; thread k: M mod q.
// out: [k + 1][kMaxLimbs].
__global__ __launch_bounds__(64) void crt_products_kernel(const uint64_t* __restrict__ primes, uint32_t k,
                                                          const uint32_t* __restrict__ q, uint32_t L,
                                                          uint32_t* __restrict__ out) {
    const uint32_t j = blockIdx.x * 64 + threadIdx.x;
    if (j > k) return;
    uint32_t qq[kMaxLimbs], x[kMaxLimbs];
    for (uint32_t l = 0; l < L; ++l) {
        qq[l] = q[l];
        x[l] = l == 0 ? 1u : 0u;
    }
    for (uint32_t i = 0; i < k; ++i)
        if (i != j) mul_word_mod(x, primes[i], qq, L);
    for (uint32_t l = 0; l < (uint32_t)kMaxLimbs; ++l) out[(size_t)j * kMaxLimbs + l] = l < L ? x[l] : 0u;
}

// One column kappa = 4 j + u of the int8 MFMA CRT table (layout, offsets and
// meaning as in nc_tables.cpp build_crt_mfma): q-limb rows hold the bytes of
// m_ju = 2^(8u) ((M / p_j) mod q) 2^32 mod q, fraction rows the bytes of
// 2^(8u) floor(2^56 / p_j), ones rows 1 at byte 0; raw bytes are stored
// offset by -128; columns of j >= k hold raw zeros.  miq: [k][kMaxLimbs].
__device__ __forceinline__ uint32_t crt_row_of(uint32_t lam, uint32_t b) {
    return 32 * (lam / 8) + 8 * ((lam % 8) / 2) + 4 * (lam % 2) + b;
}

__global__ __launch_bounds__(64) void crt_mfma_table_kernel(const uint32_t* __restrict__ miq,
                                                            const uint64_t* __restrict__ primes, uint32_t k,
                                                            const uint32_t* __restrict__ q, uint32_t L, uint32_t rt,
                                                            uint32_t kc, int8_t* __restrict__ tab) {
    const uint32_t t = blockIdx.x * 64 + threadIdx.x;
    if (t >= 32 * kc) return;
    const uint32_t j = t / 4, u = t % 4, kappa = t, kbytes = 32 * kc;
    const uint32_t lf = 8 * rt - 4, lq = lf < (uint32_t)kMaxLimbs ? lf : (uint32_t)kMaxLimbs;
    uint32_t qq[kMaxLimbs], m[kMaxLimbs];
    uint64_t g = 0;
    for (uint32_t l = 0; l < L; ++l) {
        qq[l] = q[l];
        m[l] = j < k ? miq[(size_t)j * kMaxLimbs + l] : 0u;
    }
    if (j < k) {
        for (uint32_t d = 0; d < 32 + 8 * u; ++d) dbl_mod(m, qq, L);
        g = ((uint64_t)1 << 56) / primes[j];
    }
    for (uint32_t lam = 0; lam < lq; ++lam)
        for (uint32_t b = 0; b < 4; ++b) {
            const uint32_t raw = lam < L ? (m[lam] >> (8 * b)) & 0xffu : 0u;
            tab[(size_t)crt_row_of(lam, b) * kbytes + kappa] = (int8_t)((int)raw - 128);
        }
    const uint64_t gu = g << (8 * u);
    for (uint32_t e = 0; e < 2; ++e)
        for (uint32_t b = 0; b < 4; ++b)
            tab[(size_t)crt_row_of(lf + e, b) * kbytes + kappa] =
                (int8_t)((int)((gu >> (8 * (4 * e + b))) & 0xffu) - 128);
    for (uint32_t e = 2; e < 4; ++e) tab[(size_t)crt_row_of(lf + e, 0) * kbytes + kappa] = 1;
}

// Per pseudo-limb corrections sum_b 2^(8b) (128 SA - 128^2 Kdim), SA the
// raw-byte row sums (q-limb and fraction rows), and, in thread 8 rt, the
// Montgomery-finish constant ((-M) mod q) 2^32 mod q from neg_m.
__global__ __launch_bounds__(64) void crt_mfma_corr_kernel(const int8_t* __restrict__ tab, uint32_t rt, uint32_t kc,
                                                           const uint32_t* __restrict__ neg_m,
                                                           const uint32_t* __restrict__ q, uint32_t L,
                                                           int64_t* __restrict__ corr, uint32_t* __restrict__ neg_m_r) {
    const uint32_t lam = blockIdx.x * 64 + threadIdx.x;
    const uint32_t lf = 8 * rt - 4, lq = lf < (uint32_t)kMaxLimbs ? lf : (uint32_t)kMaxLimbs;
    const uint32_t kbytes = 32 * kc;
    if (lam < 8 * rt) {
        int64_t c = 0;
        if (lam < lq || lam == lf || lam == lf + 1) {
            const int64_t off = 128ll * 128ll * kbytes;
            for (uint32_t b = 0; b < 4; ++b) {
                int64_t sa = 0;
                for (uint32_t kappa = 0; kappa < kbytes; ++kappa) sa += (int)tab[(size_t)crt_row_of(lam, b) * kbytes + kappa] + 128;
                c += (128 * sa - off) * ((int64_t)1 << (8 * b));
            }
        }
        corr[lam] = c;
    } else if (lam == 8 * rt) {
        uint32_t qq[kMaxLimbs], x[kMaxLimbs];
        for (uint32_t l = 0; l < L; ++l) {
            qq[l] = q[l];
            x[l] = neg_m[l];
        }
        for (uint32_t d = 0; d < 32; ++d) dbl_mod(x, qq, L);
        for (uint32_t l = 0; l < (uint32_t)kMaxLimbs; ++l) neg_m_r[l] = l < L ? x[l] : 0u;
    }
}

hipError_t launch_prime_search(uint64_t top, uint64_t step, uint32_t count, uint64_t* out, hipStream_t s) {
    if (count == 0 || step == 0 || top < step) return hipErrorInvalidValue;
    prime_search_kernel<<<(count + 63) / 64, 64, 0, s>>>(top, step, count, out);
    return hipGetLastError();
}

hipError_t launch_crt_products(const uint64_t* primes, uint32_t k, const uint32_t* q, uint32_t L, uint32_t* out,
                               hipStream_t s) {
    if (k == 0 || L == 0 || L > (uint32_t)kMaxLimbs) return hipErrorInvalidValue;
    crt_products_kernel<<<(k + 1 + 63) / 64, 64, 0, s>>>(primes, k, q, L, out);
    return hipGetLastError();
}

hipError_t launch_crt_mfma_tables(const uint32_t* miq, const uint64_t* primes, uint32_t k, const uint32_t* q,
                                  uint32_t L, uint32_t rt, uint32_t kc, const uint32_t* neg_m, int8_t* tab,
                                  int64_t* corr, uint32_t* neg_m_r, hipStream_t s) {
    if (k == 0 || L == 0 || L > (uint32_t)kMaxLimbs || rt == 0 || rt > (uint32_t)kCrtMaxRT || kc == 0 ||
        kc > (uint32_t)kCrtMaxKC || k > 8 * kc)
        return hipErrorInvalidValue;
    crt_mfma_table_kernel<<<(32 * kc + 63) / 64, 64, 0, s>>>(miq, primes, k, q, L, rt, kc, tab);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    crt_mfma_corr_kernel<<<(8 * rt + 1 + 63) / 64, 64, 0, s>>>(tab, rt, kc, neg_m, q, L, corr, neg_m_r);
    return hipGetLastError();
}

// ===========================================================================
// Transfers between mapped pinned host memory and device memory done by the
// GPU's own loads and stores over PCIe (the serving loop's small operands in
// and results out).  One launch moves up to kHostXferMax buffers, so a batch
// of small requests pays one launch instead of a DMA set-up per buffer
// (~13 us each, which held 240 KB copies to ~13 GB/s); 32 workgroups with
// 64 KB in flight each reach the link's ~56 GB/s in either direction while
// holding few compute-unit slots.  The host side of every access is system
// scope (sc0 sc1): coherent with the host whatever the allocation's caching,
// with nothing of it kept in the GPU's caches across launches.
// ===========================================================================
struct HostXferList {
    const uint64_t* src[kHostXferMax];
    uint64_t* dst[kHostXferMax];
};

// blockIdx.y: the buffer; a grid-stride walk over its n8 words, eight loads
// in flight per thread before their stores.
template <bool kToHost>
__global__ __launch_bounds__(256) void host_xfer_kernel(const HostXferList l, uint64_t n8) {
    const uint64_t* __restrict__ src = l.src[blockIdx.y];
    uint64_t* __restrict__ dst = l.dst[blockIdx.y];
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    auto load = [&](uint64_t i) {
        return kToHost ? src[i] : __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    };
    auto store = [&](uint64_t i, uint64_t v) {
        if (kToHost) __hip_atomic_store(dst + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        else dst[i] = v;
    };
    uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + 7 * stride < n8; i += 8 * stride) {
        uint64_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = load(i + k * stride);
#pragma unroll
        for (int k = 0; k < 8; ++k) store(i + k * stride, v[k]);
    }
    for (; i < n8; i += stride) store(i, load(i));
}

hipError_t launch_host_xfer(const void* const* src, void* const* dst, uint32_t count, uint64_t bytes, bool to_host,
                            hipStream_t s) {
    if (bytes % 8 || !src || !dst) return hipErrorInvalidValue;
    for (uint32_t i = 0; i < count; ++i)
        if (!src[i] || !dst[i] || reinterpret_cast<uintptr_t>(src[i]) % 8 || reinterpret_cast<uintptr_t>(dst[i]) % 8)
            return hipErrorInvalidValue;
    const uint64_t n8 = bytes / 8;
    if (n8 == 0 || count == 0) return hipSuccess;
    for (uint32_t off = 0; off < count; off += kHostXferMax) {
        const uint32_t c = std::min<uint32_t>(kHostXferMax, count - off);
        HostXferList l{};
        for (uint32_t i = 0; i < c; ++i) {
            l.src[i] = static_cast<const uint64_t*>(src[off + i]);
            l.dst[i] = static_cast<uint64_t*>(dst[off + i]);
        }
        // ~32 workgroups in all, none with less than 2048 words to move
        const uint64_t gx = std::max<uint64_t>(1, std::min<uint64_t>((32 + c - 1) / c, (n8 + 2047) / 2048));
        const dim3 grid((uint32_t)gx, c);
        if (to_host) host_xfer_kernel<true><<<grid, 256, 0, s>>>(l, n8);
        else host_xfer_kernel<false><<<grid, 256, 0, s>>>(l, n8);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ===========================================================================
// Scratch of nc_multiply_device_noalloc: the residue rows of a and b of an
// allocation-free call.  Each device holds its own copy for the life of the
// process; host code uses the address hipGetSymbolAddress returns for the
// current device, never the array's name.
// ===========================================================================
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

hipError_t scratch_address(void** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *out = addr[dev];
    return hipSuccess;
}

// Operands for timing runs: limb i is the high word of SplitMix64 output i
// (Steele, Lea and Flood's mixer); in each coefficient, limb hi (q's highest
// non-zero limb, of value top >= 1) is scaled into [0, top) and the limbs
// above it are zero, so every coefficient is below q.  Timing needs such
// data: with all-zero operands the multiply-bound kernels draw less power and
// run about 10 % faster than on random ones.
__global__ __launch_bounds__(256) void fill_operands_kernel(uint32_t* __restrict__ out, uint64_t n, uint32_t L,
                                                            uint32_t hi, uint32_t top, uint64_t seed) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        uint32_t v = (uint32_t)(z >> 32);
        const uint32_t l = (uint32_t)(i % L);
        if (l == hi) v = (uint32_t)(((uint64_t)v * top) >> 32);
        out[i] = l > hi ? 0u : v;
    }
}

hipError_t launch_fill_operands(uint32_t* out, uint64_t n, uint32_t L, uint32_t hi, uint32_t top, uint64_t seed,
                                hipStream_t s) {
    if (!out || L == 0 || hi >= L || top == 0) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
    fill_operands_kernel<<<(uint32_t)blocks, 256, 0, s>>>(out, n, L, hi, top, seed);
    return hipGetLastError();
}

hipError_t launch_mul32_peak(const uint32_t* in, uint64_t* out, uint32_t blocks, uint32_t iters, hipStream_t s) {
    mul32_peak_kernel<<<blocks, 256, 0, s>>>(in, out, iters);
    return hipGetLastError();
}

hipError_t launch_twiddles62(Tw* fwd, Tw* inv, const Tw* pow, const PrimeConst* pc, uint32_t log_n, uint32_t k,
                             hipStream_t s) {
    if (log_n < 1 || log_n > (uint32_t)kMaxLogN || k == 0) return hipErrorInvalidValue;
    const uint64_t n = (uint64_t)k << log_n;
    twiddles62_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(fwd, inv, pow, pc, log_n, k);
    return hipGetLastError();
}

hipError_t launch_twiddles30(Tw32* fwd, Tw32* inv, const Tw32* pow, const PrimeConst32* pc, uint32_t log_n,
                             uint32_t k, hipStream_t s) {
    if (log_n < 1 || log_n > (uint32_t)kMaxLogN || k == 0) return hipErrorInvalidValue;
    const uint64_t n = (uint64_t)k << log_n;
    twiddles30_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(fwd, inv, pow, pc, log_n, k);
    return hipGetLastError();
}

}  // namespace nc
