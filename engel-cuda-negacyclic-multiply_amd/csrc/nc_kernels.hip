// gfx950 kernels of the negacyclic product c = a*b mod (X^N + 1, q).
//
// Pipeline per product (reference: solve.cu:1489-1528):
//   decompose   limbs -> residues mod k primes        (solve.cu:533-570)
//   NTT product forward negacyclic NTT of a and b, pointwise product,
//               inverse NTT                           (solve.cu:1006-1044, 620-642)
//   CRT         residues -> L limbs mod q             (solve.cu:752-836)
//
// Arithmetic: primes p < 2^62.  64x64-bit products are built by the compiler
// from v_mad_u64_u32 / v_mul_hi_u32 (4 multiplies for a high half, 3 for a
// low half).  Twiddle products use Shoup's precomputed quotient, butterflies
// keep Harvey's lazy ranges ([0,4p) forward, [0,2p) inverse), the pointwise
// product is a Montgomery product whose 2^-64 factor is undone together with
// N^-1 in the last inverse stage.
//
// NTT formulation: negacyclic NTT with the psi twist merged into the
// butterflies (Cooley-Tukey forward, natural -> bit-reversed order;
// Gentleman-Sande inverse, bit-reversed -> natural), so neither a bit
// reversal pass nor separate twist / untwist passes exist.  Stage s (0-based)
// pairs indices that differ in bit beta = logN - 1 - s; its twiddle for the
// group i = x >> (beta + 1) is tw[2^s + i] with tw[x] = psi^bitrev(x)
// (forward) or psi^-bitrev(x) (inverse).
//
// Work decomposition of the NTT: a prime row of N values is split into
// 2^s1 contiguous blocks of 2^b values (b <= 14, so a block fits one
// workgroup's LDS).  The first s1 forward stages (bits >= b) and the last s1
// inverse stages run in "outer" kernels where each thread owns one column of
// 2^s1 values; all remaining stages run in the "inner" kernel, one block per
// workgroup, values in registers (16 per thread), LDS for the exchanges
// between radix-16 passes.  The inner kernel also does the pointwise product,
// so forward-inner, pointwise and inverse-inner share one HBM round trip.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nc_common.hpp"
#include "nc_kernels.hpp"

namespace nc {

using u128 = unsigned __int128;

// ---------------------------------------------------------------------------
// word arithmetic
// ---------------------------------------------------------------------------

// a*w mod p, lazily in [0, 2p), for any 64-bit a and w < p (Shoup).
__device__ __forceinline__ uint64_t shoup_lazy(uint64_t a, uint64_t w, uint64_t ws, uint64_t p) {
    return a * w - __umul64hi(a, ws) * p;
}

__device__ __forceinline__ uint64_t sub_if_ge(uint64_t x, uint64_t m) { return x >= m ? x - m : x; }

// Harvey CT butterfly: x, y in [0, 4p) -> (x + wy, x - wy) in [0, 4p).
__device__ __forceinline__ void ct_bfly(uint64_t& x, uint64_t& y, Tw t, uint64_t p, uint64_t p2) {
    const uint64_t u = sub_if_ge(x, p2);
    const uint64_t v = shoup_lazy(y, t.w, t.ws, p);
    x = u + v;
    y = u - v + p2;
}

// Harvey GS butterfly: x, y in [0, 2p) -> (x + y, w(x - y)) in [0, 2p).
__device__ __forceinline__ void gs_bfly(uint64_t& x, uint64_t& y, Tw t, uint64_t p, uint64_t p2) {
    const uint64_t u = x + y;
    const uint64_t d = x - y + p2;
    x = sub_if_ge(u, p2);
    y = shoup_lazy(d, t.w, t.ws, p);
}

// Last GS stage (the single group of stage 0) with N^-1 * 2^64 merged:
// x' = (x + y) nr, y' = (x - y) psi^-1 nr, both in [0, 2p).
__device__ __forceinline__ void gs_last(uint64_t& x, uint64_t& y, const PrimeConst& c) {
    const uint64_t u = x + y;
    const uint64_t d = x - y + c.p2;
    x = shoup_lazy(u, c.nr, c.nr_s, c.p);
    y = shoup_lazy(d, c.wr, c.wr_s, c.p);
}

// Montgomery product a b 2^-64 mod p for a, b < 2p, result in [0, 2p).
__device__ __forceinline__ uint64_t mont_lazy(uint64_t a, uint64_t b, uint64_t p, uint64_t mneg) {
    const u128 t = (u128)a * b;
    const uint64_t lo = (uint64_t)t, hi = (uint64_t)(t >> 64);
    const uint64_t m = lo * mneg;
    return hi + __umul64hi(m, p) + (lo != 0);
}

// Montgomery reduction of a 128-bit s < 2^64 p: s 2^-64 mod p in [0, 2p).
__device__ __forceinline__ uint64_t redc_lazy(uint64_t hi, uint64_t lo, uint64_t p, uint64_t mneg) {
    const uint64_t m = lo * mneg;
    return hi + __umul64hi(m, p) + (lo != 0);
}

// Workgroups go round-robin over the 8 XCDs (id % 8).  Give each XCD one
// contiguous range of logical ids so that neighbouring work items (same
// prime, hence same twiddle rows) share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t g, uint32_t G) {
    const uint32_t per = G >> 3;
    if (g >= (per << 3)) return g;
    return (g & 7u) * per + (g >> 3);
}

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// ---------------------------------------------------------------------------
// decomposition: limbs -> canonical residues
//   r = sum_l limb_l * (2^(32 l) 2^64 mod p)   (< 2^99, exact in 128 bits)
//   residue = REDC(r) = sum_l limb_l 2^(32 l) mod p
// A workgroup stages 64 coefficients' limbs in LDS once (coalesced) and its
// four waves cover all k primes (wave w takes primes w, w+4, ...), so each
// limb is read from HBM once for all primes.
// ---------------------------------------------------------------------------
template <int LT>
__global__ __launch_bounds__(256) void decompose_kernel(const uint32_t* __restrict__ a, uint64_t* __restrict__ res_a,
                                                        uint32_t na, const uint32_t* __restrict__ b,
                                                        uint64_t* __restrict__ res_b, const uint64_t* __restrict__ place,
                                                        const PrimeConst* __restrict__ pc, uint32_t N, uint32_t L,
                                                        uint32_t k) {
    constexpr int TILE = 64;
    constexpr int STRIDE = LT + 1;  // odd stride: conflict-free LDS reads per coefficient
    __shared__ uint32_t tile[TILE * STRIDE];
    const uint32_t tiles = (N + TILE - 1) / TILE;
    const uint32_t poly = blockIdx.x / tiles;
    const uint32_t base = (blockIdx.x % tiles) * TILE;
    const uint32_t cnt = N - base < (uint32_t)TILE ? N - base : (uint32_t)TILE;
    const bool is_a = poly < na;
    const uint32_t pi = is_a ? poly : poly - na;
    const uint32_t* src = (is_a ? a : b) + ((size_t)pi * N + base) * L;
    uint64_t* dst = (is_a ? res_a : res_b) + (size_t)pi * k * N + base;

    const uint32_t words = cnt * L;
    for (uint32_t w = threadIdx.x; w < words; w += 256) {
        const uint32_t c = w / L, l = w - c * L;
        tile[c * STRIDE + l] = src[w];
    }
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
        const uint64_t r = redc_lazy((uint64_t)(s >> 64), (uint64_t)s, p, pc[j].mont_neg);
        dst[(size_t)j * N + c] = sub_if_ge(r, p);
    }
}

// ---------------------------------------------------------------------------
// NTT passes on registers.  A pass handles R consecutive index bits
// [lo, lo + R) of the values a thread holds.  Register e = (eq << R) | ep:
// ep enumerates the pass bits, eq (when R < 4) enumerates 4 - R further bits
// that lie below the pass.  gx0 is the global row index of register 0.
// ---------------------------------------------------------------------------
template <int R, int NE>
__device__ __forceinline__ void ct_pass(uint64_t (&v)[NE], uint32_t gx0, int lo, int log_n,
                                        const Tw* __restrict__ tw, uint64_t p, uint64_t p2) {
#pragma unroll
    for (int kk = R - 1; kk >= 0; --kk) {
        const int beta = lo + kk;
        const uint32_t base = (1u << (log_n - 1 - beta)) + (gx0 >> (beta + 1));
        constexpr int MAXH = 1 << (R - 1);
        Tw t[MAXH];
        const int nh = 1 << (R - 1 - kk);
#pragma unroll
        for (int h = 0; h < MAXH; ++h)
            if (h < nh) t[h] = tw[base + h];
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int ep = e & ((1 << R) - 1);
            if (ep & (1 << kk)) continue;
            ct_bfly(v[e], v[e | (1 << kk)], t[ep >> (kk + 1)], p, p2);
        }
    }
}

template <int R, int NE>
__device__ __forceinline__ void gs_pass(uint64_t (&v)[NE], uint32_t gx0, int lo, int log_n,
                                        const Tw* __restrict__ tw, const PrimeConst& c) {
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
        const int beta = lo + kk;
        if (beta == log_n - 1) {  // stage 0: merged N^-1 2^64 scaling
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                const int ep = e & ((1 << R) - 1);
                if (ep & (1 << kk)) continue;
                gs_last(v[e], v[e | (1 << kk)], c);
            }
            continue;
        }
        const uint32_t base = (1u << (log_n - 1 - beta)) + (gx0 >> (beta + 1));
        constexpr int MAXH = 1 << (R - 1);
        Tw t[MAXH];
        const int nh = 1 << (R - 1 - kk);
#pragma unroll
        for (int h = 0; h < MAXH; ++h)
            if (h < nh) t[h] = tw[base + h];
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int ep = e & ((1 << R) - 1);
            if (ep & (1 << kk)) continue;
            gs_bfly(v[e], v[e | (1 << kk)], t[ep >> (kk + 1)], c.p, c.p2);
        }
    }
}

// ---------------------------------------------------------------------------
// inner kernel: one 2^LOGB block of one prime row per (sub-)workgroup
// ---------------------------------------------------------------------------
template <int LOGB>
struct InnerShape {
    static constexpr int TPB = 1 << (LOGB - 4);             // threads per block (16 values each)
    static constexpr int R0 = LOGB - 4 * ((LOGB - 1) / 4);  // bits of pass 0 (1..4)
    static constexpr int NP = 1 + (LOGB - R0) / 4;          // passes
    static constexpr int BPW = TPB >= 256 ? 1 : 256 / TPB;  // blocks per workgroup
    static constexpr int WG = BPW * TPB;
    static constexpr int PADB = (1 << LOGB) + (1 << (LOGB - 4));  // one pad word per 16
};

// Local index of register e of thread tid in pass m.
template <int LOGB>
__device__ __forceinline__ uint32_t xmap(int m, uint32_t tid, int e) {
    using S = InnerShape<LOGB>;
    if (m == 0) {
        const uint32_t eq = (uint32_t)e >> S::R0, ep = (uint32_t)e & ((1u << S::R0) - 1);
        return eq | (tid << (4 - S::R0)) | (ep << (LOGB - S::R0));
    }
    const int lo = LOGB - S::R0 - 4 * m;
    return (tid & ((1u << lo) - 1)) | ((uint32_t)e << lo) | ((tid >> lo) << (lo + 4));
}

__device__ __forceinline__ uint32_t lds_pad(uint32_t x) { return x + (x >> 4); }

template <int LOGB>
__device__ __forceinline__ void exchange(uint64_t (&v)[16], uint64_t* lds, uint32_t tid, int from, int to) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) lds[lds_pad(xmap<LOGB>(from, tid, e))] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = lds[lds_pad(xmap<LOGB>(to, tid, e))];
}

template <int LOGB>
__device__ __forceinline__ void forward_block(uint64_t (&v)[16], uint64_t* lds, uint32_t tid, uint32_t gblk,
                                              int log_n, const Tw* tw, uint64_t p, uint64_t p2) {
    using S = InnerShape<LOGB>;
    ct_pass<S::R0, 16>(v, gblk | xmap<LOGB>(0, tid, 0), LOGB - S::R0, log_n, tw, p, p2);
#pragma unroll
    for (int m = 1; m < S::NP; ++m) {
        exchange<LOGB>(v, lds, tid, m - 1, m);
        ct_pass<4, 16>(v, gblk | xmap<LOGB>(m, tid, 0), LOGB - S::R0 - 4 * m, log_n, tw, p, p2);
    }
}

template <int LOGB>
__device__ __forceinline__ void inverse_block(uint64_t (&v)[16], uint64_t* lds, uint32_t tid, uint32_t gblk,
                                              int log_n, const Tw* tw, const PrimeConst& c) {
    using S = InnerShape<LOGB>;
#pragma unroll
    for (int m = S::NP - 1; m >= 1; --m) {
        gs_pass<4, 16>(v, gblk | xmap<LOGB>(m, tid, 0), LOGB - S::R0 - 4 * m, log_n, tw, c);
        exchange<LOGB>(v, lds, tid, m, m - 1);
    }
    gs_pass<S::R0, 16>(v, gblk | xmap<LOGB>(0, tid, 0), LOGB - S::R0, log_n, tw, c);
}

// ra, rb: [nprod][k][N] residue rows after the outer forward stages (or
// canonical residues when the block is the whole row).  Output overwrites ra:
// lazy [0, 2p) values awaiting the outer inverse stages, or canonical
// residues of the product when the block is the whole row.
template <int LOGB>
__global__ __launch_bounds__(InnerShape<LOGB>::WG) void ntt_inner_kernel(
    uint64_t* __restrict__ ra, const uint64_t* __restrict__ rb, const Tw* __restrict__ twf,
    const Tw* __restrict__ twi, const PrimeConst* __restrict__ pcs, uint32_t log_n, uint32_t k, uint32_t nprod) {
    using S = InnerShape<LOGB>;
    __shared__ uint64_t lds_all[S::BPW * S::PADB];
    const uint32_t sub = threadIdx.x / S::TPB;
    const uint32_t tid = threadIdx.x % S::TPB;
    uint64_t* lds = lds_all + sub * S::PADB;
    const uint32_t N = 1u << log_n;
    const uint32_t nblk_log = log_n - LOGB;
    const uint32_t total = k * nprod << nblk_log;
    uint32_t wid = xcd_remap(blockIdx.x, gridDim.x) * S::BPW + sub;
    const bool active = wid < total;
    if (!active) wid = 0;  // idle sub-block: keeps barriers uniform, touches nothing
    const uint32_t blk = wid & ((1u << nblk_log) - 1);
    const uint32_t row = wid >> nblk_log;  // = j * nprod + prod
    const uint32_t j = row / nprod, prod = row - j * nprod;
    const size_t off = ((size_t)prod * k + j) * N + ((size_t)blk << LOGB);
    const uint32_t gblk = blk << LOGB;
    const PrimeConst c = pcs[j];
    const Tw* tf = twf + (size_t)j * N;
    const Tw* ti = twi + (size_t)j * N;
    const bool final_out = (nblk_log == 0);

    uint64_t va[16], vb[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) va[e] = active ? ra[off + xmap<LOGB>(0, tid, e)] : 0;
#pragma unroll
    for (int e = 0; e < 16; ++e) vb[e] = active ? rb[off + xmap<LOGB>(0, tid, e)] : 0;
    forward_block<LOGB>(va, lds, tid, gblk, (int)log_n, tf, c.p, c.p2);
    forward_block<LOGB>(vb, lds, tid, gblk, (int)log_n, tf, c.p, c.p2);
#pragma unroll
    for (int e = 0; e < 16; ++e) va[e] = mont_lazy(sub_if_ge(va[e], c.p2), sub_if_ge(vb[e], c.p2), c.p, c.mont_neg);
    inverse_block<LOGB>(va, lds, tid, gblk, (int)log_n, ti, c);
    if (!active) return;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const uint64_t x = final_out ? sub_if_ge(va[e], c.p) : va[e];
        ra[off + xmap<LOGB>(0, tid, e)] = x;
    }
}

// ---------------------------------------------------------------------------
// outer kernels: the first S1 forward stages / last S1 inverse stages.
// Thread = one column x = col + (e << b), e in [0, 2^S1), b = logN - S1.
// ---------------------------------------------------------------------------
template <int S1>
__global__ __launch_bounds__(256) void ntt_outer_fwd_kernel(uint64_t* __restrict__ ra, uint64_t* __restrict__ rb,
                                                            const Tw* __restrict__ twf,
                                                            const PrimeConst* __restrict__ pcs, uint32_t log_n,
                                                            uint32_t k, uint32_t nprod) {
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
    uint64_t* base = (which ? rb : ra) + ((size_t)prod * k + j) * ((size_t)1 << log_n) + col;
    const PrimeConst c = pcs[j];
    uint64_t v[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) v[e] = base[(size_t)e << b];
    ct_pass<S1, NE>(v, col, (int)b, (int)log_n, twf + ((size_t)j << log_n), c.p, c.p2);
#pragma unroll
    for (int e = 0; e < NE; ++e) base[(size_t)e << b] = v[e];
}

template <int S1>
__global__ __launch_bounds__(256) void ntt_outer_inv_kernel(uint64_t* __restrict__ ra, const Tw* __restrict__ twi,
                                                            const PrimeConst* __restrict__ pcs, uint32_t log_n,
                                                            uint32_t k, uint32_t nprod) {
    constexpr int NE = 1 << S1;
    const uint32_t b = log_n - S1;
    const uint32_t cols = 1u << b;
    const uint32_t wgs_per_row = (cols + 255) / 256;
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t row = g / wgs_per_row;  // (j * nprod + prod)
    const uint32_t col = (g - row * wgs_per_row) * 256 + threadIdx.x;
    if (col >= cols) return;
    const uint32_t j = row / nprod, prod = row - j * nprod;
    uint64_t* base = ra + ((size_t)prod * k + j) * ((size_t)1 << log_n) + col;
    const PrimeConst c = pcs[j];
    uint64_t v[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) v[e] = base[(size_t)e << b];
    gs_pass<S1, NE>(v, col, (int)b, (int)log_n, twi + ((size_t)j << log_n), c);
#pragma unroll
    for (int e = 0; e < NE; ++e) base[(size_t)e << b] = sub_if_ge(v[e], c.p);
}

// ---------------------------------------------------------------------------
// tiny rows (N < 16): one thread per (product, prime), plain loops.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void ntt_tiny_kernel(uint64_t* __restrict__ ra, const uint64_t* __restrict__ rb,
                                                      const Tw* __restrict__ twf, const Tw* __restrict__ twi,
                                                      const PrimeConst* __restrict__ pcs, uint32_t log_n, uint32_t k,
                                                      uint32_t nprod) {
    const uint32_t t = blockIdx.x * 64 + threadIdx.x;
    if (t >= k * nprod) return;
    const uint32_t j = t / nprod, prod = t - j * nprod;
    const uint32_t N = 1u << log_n;
    const size_t off = ((size_t)prod * k + j) * N;
    const PrimeConst c = pcs[j];
    const Tw* tf = twf + (size_t)j * N;
    const Tw* ti = twi + (size_t)j * N;
    uint64_t x[8], y[8];
    for (uint32_t i = 0; i < N; ++i) { x[i] = ra[off + i]; y[i] = rb[off + i]; }
    for (int s = 0; s < (int)log_n; ++s) {
        const int beta = (int)log_n - 1 - s;
        for (uint32_t i = 0; i < N; ++i) {
            if (i & (1u << beta)) continue;
            const Tw w = tf[(1u << s) + (i >> (beta + 1))];
            ct_bfly(x[i], x[i | (1u << beta)], w, c.p, c.p2);
            ct_bfly(y[i], y[i | (1u << beta)], w, c.p, c.p2);
        }
    }
    for (uint32_t i = 0; i < N; ++i) x[i] = mont_lazy(sub_if_ge(x[i], c.p2), sub_if_ge(y[i], c.p2), c.p, c.mont_neg);
    for (int beta = 0; beta < (int)log_n; ++beta) {
        const int s = (int)log_n - 1 - beta;
        for (uint32_t i = 0; i < N; ++i) {
            if (i & (1u << beta)) continue;
            if (s == 0) gs_last(x[i], x[i | (1u << beta)], c);
            else gs_bfly(x[i], x[i | (1u << beta)], ti[(1u << s) + (i >> (beta + 1))], c.p, c.p2);
        }
    }
    for (uint32_t i = 0; i < N; ++i) ra[off + i] = sub_if_ge(x[i], c.p);
}

// ---------------------------------------------------------------------------
// CRT: canonical residues -> L limbs mod q.
//   v_j   = r_j (M/p_j)^-1 mod p_j                   (Shoup, exact)
//   acc   = sum_j v_j ((M/p_j) mod q)                (LT + 3 limbs)
//   frac  = sum_j floor(v_j floor(2^128/p_j) / 2^64) (128-bit fixed point)
//   K     = floor(frac / 2^64) + [frac mod 2^64 >= 2^63]
//   c     = (acc + K ((-M) mod q)) mod q
// K counts the multiples of M in sum_j v_j M/p_j plus the centred-lift
// correction, exactly as solve.cu:779-834 subtracts sum_hi (M mod q) and,
// when the fraction is >= 1/2, M mod q once more; folding both into one
// non-negative addition leaves a single Barrett reduction.
// ---------------------------------------------------------------------------
template <int NA>
__device__ __forceinline__ void mac_limbs(uint32_t (&acc)[NA], const uint32_t* __restrict__ m, int mlen_ct,
                                          uint32_t f, int off) {
    uint64_t carry = 0;
#pragma unroll
    for (int l = 0; l < NA; ++l) {
        if (l < off) continue;
        if (l - off < mlen_ct) {
            const uint64_t t = (uint64_t)f * m[l - off] + acc[l] + carry;
            acc[l] = (uint32_t)t;
            carry = t >> 32;
        } else {
            const uint64_t t = (uint64_t)acc[l] + carry;
            acc[l] = (uint32_t)t;
            carry = t >> 32;
        }
    }
}

// r = acc mod q for acc < 2^(bits(q) + 71), by a Barrett quotient estimate
// qhat = floor((acc >> (bits(q)-1)) * mu / 2^72), mu = floor(2^(bits(q)-1+72)/q),
// which is at most 2 below the true quotient; then <= 2 subtractions of q.
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
    // t * mu, 6 limbs
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
    // qhat = pr >> 72: limbs 2.. shifted by 8 bits
    uint32_t qh[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) qh[i] = (pr[2 + i] >> 8) | (i + 3 < 6 ? (pr[3 + i] << 24) : 0u);
    // acc -= qhat * q
    uint32_t prod[NA];
#pragma unroll
    for (int l = 0; l < NA; ++l) prod[l] = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) mac_limbs<NA>(prod, qc->q, LT, qh[i], i);
    uint32_t borrow = 0;
#pragma unroll
    for (int l = 0; l < NA; ++l) {
        const uint64_t d = (uint64_t)acc[l] - prod[l] - borrow;
        acc[l] = (uint32_t)d;
        borrow = (uint32_t)(d >> 63);
    }
    // at most two subtractions of q
#pragma unroll
    for (int fix = 0; fix < 2; ++fix) {
        uint32_t diff[NA];
        uint32_t br = 0;
#pragma unroll
        for (int l = 0; l < NA; ++l) {
            const uint64_t d = (uint64_t)acc[l] - (l < LT ? qc->q[l] : 0u) - br;
            diff[l] = (uint32_t)d;
            br = (uint32_t)(d >> 63);
        }
        if (!br) {
#pragma unroll
            for (int l = 0; l < NA; ++l) acc[l] = diff[l];
        }
    }
}

template <int LT>
__global__ __launch_bounds__(256) void crt_kernel(const uint64_t* __restrict__ res, uint32_t* __restrict__ out,
                                                  const CrtConst* __restrict__ cc, const uint32_t* __restrict__ miq,
                                                  const QConst* __restrict__ qc, const PrimeConst* __restrict__ pcs,
                                                  uint32_t log_n, uint32_t L, uint32_t k, uint32_t nprod) {
    constexpr int NA = LT + 3;
    const uint32_t N = 1u << log_n;
    const uint32_t idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= (nprod << log_n)) return;
    const uint32_t prod = idx >> log_n, i = idx & (N - 1);
    const uint64_t* r = res + (size_t)prod * k * N + i;
    uint32_t acc[NA];
#pragma unroll
    for (int l = 0; l < NA; ++l) acc[l] = 0;
    uint64_t f_lo = 0, f_hi = 0;
    for (uint32_t j = 0; j < k; ++j) {
        const uint64_t p = pcs[j].p;
        const CrtConst cj = cc[j];
        const uint64_t v = sub_if_ge(shoup_lazy(r[(size_t)j * N], cj.y, cj.y_s, p), p);
        const uint64_t f = v * cj.mag_hi + __umul64hi(v, cj.mag_lo);
        f_lo += f;
        f_hi += (f_lo < f);
        const uint32_t* m = miq + (size_t)j * kMaxLimbs;
        mac_limbs<NA>(acc, m, LT, (uint32_t)v, 0);
        mac_limbs<NA>(acc, m, LT, (uint32_t)(v >> 32), 1);
    }
    const uint32_t K = (uint32_t)f_hi + (uint32_t)(f_lo >> 63);
    mac_limbs<NA>(acc, qc->neg_m, LT, K, 0);
    barrett_mod_q<LT>(acc, qc);
    uint32_t* o = out + ((size_t)prod * N + i) * L;
#pragma unroll
    for (int l = 0; l < LT; ++l)
        if ((uint32_t)l < L) o[l] = acc[l];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
uint32_t max_log_blocks(uint32_t log_n) {
    if (log_n < 4) return 0;
    const uint32_t m = log_n - 4;  // block must keep >= 16 values
    return m < (uint32_t)kMaxOuterBits ? m : (uint32_t)kMaxOuterBits;
}

uint32_t min_log_blocks(uint32_t log_n) { return log_n > (uint32_t)kMaxLogBlock ? log_n - kMaxLogBlock : 0u; }

hipError_t launch_decompose(const DevPlan& d, const uint32_t* a, uint64_t* res_a, uint32_t na, const uint32_t* b,
                            uint64_t* res_b, uint32_t nb, hipStream_t s) {
    const uint32_t tiles = (d.N + 63) / 64;
    const dim3 grid(tiles * (na + nb));
    if (d.L == 28)
        decompose_kernel<28><<<grid, 256, 0, s>>>(a, res_a, na, b, res_b, d.place, d.pc, d.N, d.L, d.k);
    else
        decompose_kernel<kMaxLimbs><<<grid, 256, 0, s>>>(a, res_a, na, b, res_b, d.place, d.pc, d.N, d.L, d.k);
    return hipGetLastError();
}

template <int LOGB>
static void launch_inner(const DevPlan& d, uint64_t* ra, const uint64_t* rb, uint32_t nprod, hipStream_t s) {
    using S = InnerShape<LOGB>;
    const uint32_t items = (d.k * nprod) << (d.log_n - LOGB);
    const uint32_t grid = (items + S::BPW - 1) / S::BPW;
    ntt_inner_kernel<LOGB><<<grid, S::WG, 0, s>>>(ra, rb, d.tw_fwd, d.tw_inv, d.pc, d.log_n, d.k, nprod);
}

template <int S1>
static void launch_outer_fwd(const DevPlan& d, uint64_t* ra, uint64_t* rb, uint32_t nprod, hipStream_t s) {
    const uint32_t cols = d.N >> S1;
    const uint32_t wgs_per_row = (cols + 255) / 256;
    const uint32_t grid = wgs_per_row * 2 * d.k * nprod;
    ntt_outer_fwd_kernel<S1><<<grid, 256, 0, s>>>(ra, rb, d.tw_fwd, d.pc, d.log_n, d.k, nprod);
}

template <int S1>
static void launch_outer_inv(const DevPlan& d, uint64_t* ra, uint32_t nprod, hipStream_t s) {
    const uint32_t cols = d.N >> S1;
    const uint32_t wgs_per_row = (cols + 255) / 256;
    const uint32_t grid = wgs_per_row * d.k * nprod;
    ntt_outer_inv_kernel<S1><<<grid, 256, 0, s>>>(ra, d.tw_inv, d.pc, d.log_n, d.k, nprod);
}

hipError_t launch_ntt_outer_fwd(const DevPlan& d, uint64_t* ra, uint64_t* rb, uint32_t nprod, uint32_t log_blocks,
                                hipStream_t s) {
    if (d.log_n < 4) return hipSuccess;
    switch (log_blocks) {
        case 0: break;
        case 1: launch_outer_fwd<1>(d, ra, rb, nprod, s); break;
        case 2: launch_outer_fwd<2>(d, ra, rb, nprod, s); break;
        case 3: launch_outer_fwd<3>(d, ra, rb, nprod, s); break;
        case 4: launch_outer_fwd<4>(d, ra, rb, nprod, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ntt_inner(const DevPlan& d, uint64_t* ra, const uint64_t* rb, uint32_t nprod, uint32_t log_blocks,
                            hipStream_t s) {
    if (d.log_n < 4) {
        const uint32_t t = d.k * nprod;
        ntt_tiny_kernel<<<(t + 63) / 64, 64, 0, s>>>(ra, rb, d.tw_fwd, d.tw_inv, d.pc, d.log_n, d.k, nprod);
        return hipGetLastError();
    }
    switch (d.log_n - log_blocks) {
        case 4: launch_inner<4>(d, ra, rb, nprod, s); break;
        case 5: launch_inner<5>(d, ra, rb, nprod, s); break;
        case 6: launch_inner<6>(d, ra, rb, nprod, s); break;
        case 7: launch_inner<7>(d, ra, rb, nprod, s); break;
        case 8: launch_inner<8>(d, ra, rb, nprod, s); break;
        case 9: launch_inner<9>(d, ra, rb, nprod, s); break;
        case 10: launch_inner<10>(d, ra, rb, nprod, s); break;
        case 11: launch_inner<11>(d, ra, rb, nprod, s); break;
        case 12: launch_inner<12>(d, ra, rb, nprod, s); break;
        case 13: launch_inner<13>(d, ra, rb, nprod, s); break;
        case 14: launch_inner<14>(d, ra, rb, nprod, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ntt_outer_inv(const DevPlan& d, uint64_t* ra, uint32_t nprod, uint32_t log_blocks, hipStream_t s) {
    if (d.log_n < 4) return hipSuccess;
    switch (log_blocks) {
        case 0: break;
        case 1: launch_outer_inv<1>(d, ra, nprod, s); break;
        case 2: launch_outer_inv<2>(d, ra, nprod, s); break;
        case 3: launch_outer_inv<3>(d, ra, nprod, s); break;
        case 4: launch_outer_inv<4>(d, ra, nprod, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ntt_product(const DevPlan& d, uint64_t* ra, uint64_t* rb, uint32_t nprod, uint32_t log_blocks,
                              hipStream_t s) {
    if (d.log_n >= 4 && (log_blocks < min_log_blocks(d.log_n) || log_blocks > max_log_blocks(d.log_n)))
        return hipErrorInvalidValue;
    hipError_t e = launch_ntt_outer_fwd(d, ra, rb, nprod, log_blocks, s);
    if (e != hipSuccess) return e;
    e = launch_ntt_inner(d, ra, rb, nprod, log_blocks, s);
    if (e != hipSuccess) return e;
    return launch_ntt_outer_inv(d, ra, nprod, log_blocks, s);
}

hipError_t launch_crt(const DevPlan& d, const uint64_t* res, uint32_t* out, uint32_t nprod, uint32_t split,
                      hipStream_t s) {
    (void)split;
    const uint32_t total = nprod << d.log_n;
    const dim3 grid((total + 255) / 256);
    if (d.L == 28)
        crt_kernel<28><<<grid, 256, 0, s>>>(res, out, d.cc, d.miq, d.qc, d.pc, d.log_n, d.L, d.k, nprod);
    else
        crt_kernel<kMaxLimbs><<<grid, 256, 0, s>>>(res, out, d.cc, d.miq, d.qc, d.pc, d.log_n, d.L, d.k, nprod);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// 32x32->64 multiply throughput probe: eight independent v_mad_u64_u32 chains
// per lane, no other VALU work in the loop.  Used to measure the integer
// multiply peak that bounds this library's kernels.
// ---------------------------------------------------------------------------
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

hipError_t launch_mul32_peak(const uint32_t* in, uint64_t* out, uint32_t blocks, uint32_t iters, hipStream_t s) {
    mul32_peak_kernel<<<blocks, 256, 0, s>>>(in, out, iters);
    return hipGetLastError();
}

}  // namespace nc
