// Plan tables: the prime set and every per-prime / per-q constant the
// kernels read.  Runs once per plan on the host (the reference does the same
// work in fherma_init, solve.cu:1174-1398, outside the timed call).
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>

#include "nc_host_math.hpp"
#include "nc_plan.hpp"

namespace nc {

using namespace host;

static uint32_t bitrev(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

// Smallest base g >= 2 whose power g^((p-1)/2N) is a primitive 2N-th root
// (psi^N = -1); the same choice rule as solve.cu:495-502.
static uint64_t find_psi(uint64_t p, uint32_t N) {
    const uint64_t e = (p - 1) / (2ull * N);
    for (uint64_t g = 2; g < 100000; ++g) {
        const uint64_t psi = powmod(g, e, p);
        if (powmod(psi, N, p) == p - 1) return psi;
    }
    throw std::runtime_error("no primitive 2N-th root of unity for prime " + std::to_string(p));
}

HostTables build_tables(uint32_t N, uint32_t L, const uint32_t* q) {
    // validation (solve.cu:1167-1172; N cap raised to 65536)
    if (N < 2 || N > (1u << kMaxLogN) || (N & (N - 1))) throw std::invalid_argument("N must be a power of two from 2 to 65536");
    if (L == 0 || L > (uint32_t)kMaxLimbs) throw std::invalid_argument("L must be in [1, 32]");
    if (!q || !(q[0] & 1u)) throw std::invalid_argument("q must be an odd modulus of L limbs");
    HostTables t;
    t.N = N;
    t.L = L;
    while ((1u << t.log_n) < N) ++t.log_n;
    Big qb(L);
    for (uint32_t l = 0; l < L; ++l) qb.v[l] = q[l];
    t.q_bits = (uint32_t)qb.bits();
    if (t.q_bits < 2) throw std::invalid_argument("q must be an odd modulus > 1");
    // fast-CRT width (solve.cu:1339-1341): bits(q) + 70 <= 34 * 32
    if (t.q_bits + 70 > 1088) throw std::invalid_argument("q wider than 1018 bits is not supported");

    // prime rule (solve.cu:1179-1204): p = 1 mod 2N, descending from 2^62 - 1,
    // until bits(prod p) >= log2 N + 2 bits(q) + 4
    const uint32_t bound_bits = t.log_n + 2 * t.q_bits + 4;
    const uint64_t step = 2ull * N;
    uint64_t cand = (1ull << 62) - 1;
    cand -= (cand - 1) % step;
    const size_t mlimbs = 80;
    Big M = Big::from_u64(1, mlimbs);
    while (cand > step && (uint32_t)M.bits() < bound_bits) {
        if (is_prime(cand)) {
            if (t.primes.size() >= (size_t)kMaxPrimes) throw std::runtime_error("the coefficient bound does not fit the residue budget");
            M.mul_u64(cand);
            t.primes.push_back(cand);
        }
        cand -= step;
    }
    if ((uint32_t)M.bits() < bound_bits) throw std::runtime_error("not enough NTT primes for this modulus");
    const uint32_t k = (uint32_t)t.primes.size();
    t.k = k;

    t.tw_fwd.resize((size_t)k * N);
    t.tw_inv.resize((size_t)k * N);
    t.pc.resize(k);
    t.place.assign((size_t)k * kMaxLimbs, 0);
    t.cc.resize(k);
    t.miq.assign((size_t)k * kMaxLimbs, 0);

    std::vector<uint64_t> pw(N), ipw(N);
    for (uint32_t j = 0; j < k; ++j) {
        const uint64_t p = t.primes[j];
        const uint64_t psi = find_psi(p, N);
        const uint64_t psi_inv = invmod(psi, p);
        pw[0] = ipw[0] = 1;
        for (uint32_t i = 1; i < N; ++i) {
            pw[i] = mulmod(pw[i - 1], psi, p);
            ipw[i] = mulmod(ipw[i - 1], psi_inv, p);
        }
        Tw* tf = &t.tw_fwd[(size_t)j * N];
        Tw* ti = &t.tw_inv[(size_t)j * N];
        tf[0] = ti[0] = Tw{0, 0};
        for (uint32_t x = 1; x < N; ++x) {
            const uint32_t r = bitrev(x, t.log_n);
            tf[x] = Tw{pw[r], shoup(pw[r], p)};
            ti[x] = Tw{ipw[r], shoup(ipw[r], p)};
        }
        const uint64_t r64 = (uint64_t)(((u128)1 << 64) % p);  // 2^64 mod p
        const uint64_t ninv = invmod(N % p, p);
        PrimeConst& c = t.pc[j];
        c.p = p;
        c.p2 = 2 * p;
        c.mont_neg = mont_neg_inv(p);
        c.nr = mulmod(ninv, r64, p);
        c.nr_s = shoup(c.nr, p);
        const uint64_t last = N > 1 ? ti[1].w : 1;  // psi^-bitrev(1): the stage-0 inverse twiddle
        c.wr = mulmod(last, c.nr, p);
        c.wr_s = shoup(c.wr, p);
        c.pad = 0;
        // decomposition places 2^(32 l) 2^64 mod p (the 2^64 cancels in REDC)
        uint64_t pl = r64;
        const uint64_t two32 = (1ull << 32) % p;
        for (uint32_t l = 0; l < L; ++l) {
            t.place[(size_t)j * kMaxLimbs + l] = pl;
            pl = mulmod(pl, two32, p);
        }
    }

    // CRT constants (fast CRT, solve.cu:1342-1369 for the same quantities)
    for (uint32_t j = 0; j < k; ++j) {
        const uint64_t p = t.primes[j];
        Big Mj = Big::from_u64(1, mlimbs);
        uint64_t mj_mod_p = 1;
        for (uint32_t o = 0; o < k; ++o) {
            if (o == j) continue;
            Mj.mul_u64(t.primes[o]);
            mj_mod_p = mulmod(mj_mod_p, t.primes[o] % p, p);
        }
        CrtConst& cc = t.cc[j];
        cc.y = invmod(mj_mod_p, p);
        cc.y_s = shoup(cc.y, p);
        const u128 mag = (~(u128)0) / p;  // floor((2^128 - 1)/p) = floor(2^128/p), p odd
        cc.mag_hi = (uint64_t)(mag >> 64);
        cc.mag_lo = (uint64_t)mag;
        Big r = mod(Mj, qb);
        for (uint32_t l = 0; l < L; ++l) t.miq[(size_t)j * kMaxLimbs + l] = r.v[l];
    }
    QConst& qc = t.qc;
    std::memset(&qc, 0, sizeof qc);
    for (uint32_t l = 0; l < L; ++l) qc.q[l] = q[l];
    Big mq = mod(M, qb);
    Big neg(L);
    bool zero = mq.bits() == 0;
    if (!zero) {
        neg = qb;
        sub_in(neg, mq);
    }
    for (uint32_t l = 0; l < L; ++l) qc.neg_m[l] = neg.v[l];
    Big mu = pow2_div((int)t.q_bits - 1 + 72, qb, 4);
    for (int i = 0; i < 4; ++i) qc.mu[i] = mu.v[i];
    if (qc.mu[3] != 0) throw std::runtime_error("Barrett constant wider than 96 bits");
    qc.q_bits = t.q_bits;
    qc.L = L;
    return t;
}

// Known-answer checks of this formulation, run at plan creation.
void self_test(const HostTables& t) {
    std::mt19937_64 rng(0x5eed);
    for (uint32_t j = 0; j < t.k; ++j) {
        const PrimeConst& c = t.pc[j];
        const uint64_t p = c.p;
        if (!is_prime(p) || (p - 1) % (2ull * t.N) != 0 || p >= (1ull << 62))
            throw std::runtime_error("self test: prime rule violated");
        if ((uint64_t)(p * (0 - c.mont_neg)) != 1) throw std::runtime_error("self test: Montgomery constant");
        // psi is a primitive 2N-th root: tw_fwd[1] = psi^(N/2) squares to -1
        if (t.N > 1) {
            const uint64_t h = t.tw_fwd[(size_t)j * t.N + 1].w;
            if (mulmod(h, h, p) != p - 1) throw std::runtime_error("self test: psi order");
            if (mulmod(h, t.tw_inv[(size_t)j * t.N + 1].w, p) != 1) throw std::runtime_error("self test: psi inverse");
        }
        // Shoup lazy product lands in [0, 2p) and agrees with the exact product
        for (int trial = 0; trial < 16; ++trial) {
            const uint64_t a = rng(), w = rng() % p;
            const uint64_t ws = shoup(w, p);
            const uint64_t lazy = a * w - (uint64_t)(((u128)a * ws) >> 64) * p;
            if (lazy >= 2 * p || lazy % p != mulmod(a % p, w, p)) throw std::runtime_error("self test: Shoup product");
        }
        // place table: REDC(sum limb 2^(32l) 2^64) equals Horner over the limbs
        uint32_t limbs[kMaxLimbs];
        for (uint32_t l = 0; l < t.L; ++l) limbs[l] = (uint32_t)rng();
        u128 s = 0;
        for (uint32_t l = 0; l < t.L; ++l) s += (u128)limbs[l] * t.place[(size_t)j * kMaxLimbs + l];
        const uint64_t lo = (uint64_t)s, hi = (uint64_t)(s >> 64);
        const uint64_t m = lo * c.mont_neg;
        uint64_t redc = hi + (uint64_t)(((u128)m * p) >> 64) + (lo != 0);
        if (redc >= p) redc -= p;
        uint64_t horner = 0;
        for (int l = (int)t.L - 1; l >= 0; --l) horner = (uint64_t)((((u128)horner) << 32 | limbs[l]) % p);
        if (redc != horner) throw std::runtime_error("self test: digit sum does not match the modulus");
    }
}

}  // namespace nc
