// Plan tables: the prime sets and every per-prime / per-q constant the
// kernels read.  Runs once per plan on the host (the reference does the same
// work in fherma_init, solve.cu:1174-1398, outside the timed call).
#include <algorithm>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>

#include "nc_host_math.hpp"
#include "nc_plan.hpp"

namespace nc {

using namespace host;

namespace {

uint32_t bitrev(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

// Smallest base g >= 2 whose power g^((p-1)/2N) is a primitive 2N-th root
// (psi^N = -1); the same choice rule as solve.cu:495-502.
uint64_t find_psi(uint64_t p, uint32_t N) {
    const uint64_t e = (p - 1) / (2ull * N);
    for (uint64_t g = 2; g < 100000; ++g) {
        const uint64_t psi = powmod(g, e, p);
        if (powmod(psi, N, p) == p - 1) return psi;
    }
    throw std::runtime_error("no primitive 2N-th root of unity for prime " + std::to_string(p));
}

// Prime rule of solve.cu:1179-1204, parameterised by the word: candidates
// p = 1 mod 2N descending from 2^bits - 1, taken while prime, until
// bits(prod p) >= log2 N + 2 bits(q) + 4.
std::vector<uint64_t> pick_primes(uint32_t N, uint32_t log_n, uint32_t q_bits, int bits, size_t max_k, Big& M) {
    const uint32_t bound_bits = log_n + 2 * q_bits + 4;
    const uint64_t step = 2ull * N;
    uint64_t cand = (1ull << bits) - 1;
    cand -= (cand - 1) % step;
    M = Big::from_u64(1, 80);
    std::vector<uint64_t> primes;
    while (cand > step && (uint32_t)M.bits() < bound_bits) {
        if (is_prime(cand)) {
            if (primes.size() >= max_k) throw std::runtime_error("the coefficient bound does not fit the residue budget");
            M.mul_u64(cand);
            primes.push_back(cand);
        }
        cand -= step;
    }
    if ((uint32_t)M.bits() < bound_bits) throw std::runtime_error("not enough NTT primes for this modulus");
    return primes;
}

// Twiddles psi^bitrev(x) / psi^-bitrev(x), x in [1, N), and the per-prime
// values shared by both widths (as 64-bit numbers; the caller narrows).
struct RootTables {
    std::vector<uint64_t> fwd, inv;  // [N] (full tables only)
    std::vector<uint64_t> pf, pi;    // [log_n]: psi^(2^i), psi^-(2^i)
    uint64_t nr = 0, wr = 0;         // N^-1 R mod p, psi^-bitrev(1) N^-1 R mod p
};

RootTables roots(uint64_t p, uint32_t N, uint32_t log_n, int word_bits, bool full, uint64_t psi_given = 0) {
    const uint64_t psi = psi_given ? psi_given : find_psi(p, N);
    const uint64_t psi_inv = invmod(psi, p);
    RootTables r;
    r.pf.resize(log_n);
    r.pi.resize(log_n);
    for (uint32_t i = 0; i < log_n; ++i) {
        r.pf[i] = i ? mulmod(r.pf[i - 1], r.pf[i - 1], p) : psi;
        r.pi[i] = i ? mulmod(r.pi[i - 1], r.pi[i - 1], p) : psi_inv;
    }
    if (full) {
        std::vector<uint64_t> pw(N), ipw(N);
        pw[0] = ipw[0] = 1;
        for (uint32_t i = 1; i < N; ++i) {
            pw[i] = mulmod(pw[i - 1], psi, p);
            ipw[i] = mulmod(ipw[i - 1], psi_inv, p);
        }
        r.fwd.assign(N, 0);
        r.inv.assign(N, 0);
        for (uint32_t x = 1; x < N; ++x) {
            const uint32_t e = bitrev(x, log_n);
            r.fwd[x] = pw[e];
            r.inv[x] = ipw[e];
        }
    }
    const uint64_t R = (uint64_t)((((u128)1) << word_bits) % p);  // the Montgomery radix mod p
    r.nr = mulmod(invmod(N % p, p), R, p);
    // psi^-bitrev(1) = psi^-(N/2)
    r.wr = mulmod(N > 1 ? r.pi[log_n - 1] : 1, r.nr, p);
    return r;
}

// psi^(+-e) mod p from the squarings table, one factor per set bit of e
uint64_t pow_from_table(const std::vector<uint64_t>& sq, uint32_t e, uint64_t p) {
    uint64_t w = 1;
    for (uint32_t i = 0; e; ++i, e >>= 1)
        if (e & 1u) w = mulmod(w, sq[i], p);
    return w;
}

uint64_t shoup_w(uint64_t w, uint64_t p, int word_bits) { return (uint64_t)(((u128)w << word_bits) / p); }

uint64_t neg_inv(uint64_t p, int word_bits) {
    const uint64_t x = mont_neg_inv(p);  // -p^-1 mod 2^64
    return word_bits == 64 ? x : (x & ((1ull << word_bits) - 1));
}

// CRT constants common to both widths: miq = (M/p_j) mod q, y_j, and the
// QConst (q, (-M) mod q, Barrett constant).
// (M / p_j) mod q by host big-integer arithmetic.
Big product_mod_q(const std::vector<uint64_t>& primes, size_t skip, const Big& qb) {
    Big Mj = Big::from_u64(1, 80);
    for (size_t o = 0; o < primes.size(); ++o)
        if (o != skip) Mj.mul_u64(primes[o]);
    return mod(Mj, qb);
}

// prod (device-made, may be null): [k + 1][kMaxLimbs], (M / p_j) mod q then
// M mod q; the first and last products are recomputed here and must agree.
void crt_common(const std::vector<uint64_t>& primes, const Big& M, const Big& qb, uint32_t L, uint32_t q_bits,
                std::vector<uint32_t>& miq, std::vector<uint64_t>& y, QConst& qc,
                const std::vector<uint32_t>* prod = nullptr) {
    const size_t k = primes.size();
    miq.assign(k * kMaxLimbs, 0);
    y.assign(k, 0);
    if (prod && prod->size() != (k + 1) * kMaxLimbs) throw std::runtime_error("self test: device CRT products shape");
    for (size_t j = 0; j < k; ++j) {
        const uint64_t p = primes[j];
        uint64_t mj_mod_p = 1;
        for (size_t o = 0; o < k; ++o)
            if (o != j) mj_mod_p = mulmod(mj_mod_p, primes[o] % p, p);
        y[j] = invmod(mj_mod_p, p);
        if (prod) {
            for (uint32_t l = 0; l < L; ++l) miq[j * kMaxLimbs + l] = (*prod)[j * kMaxLimbs + l];
        } else {
            Big r = product_mod_q(primes, j, qb);
            for (uint32_t l = 0; l < L; ++l) miq[j * kMaxLimbs + l] = r.v[l];
        }
    }
    if (prod) {
        for (size_t j : {(size_t)0, k - 1}) {
            Big r = product_mod_q(primes, j, qb);
            for (uint32_t l = 0; l < L; ++l)
                if (miq[j * kMaxLimbs + l] != r.v[l]) throw std::runtime_error("self test: device CRT products");
        }
    }
    std::memset(&qc, 0, sizeof qc);
    for (uint32_t l = 0; l < L; ++l) qc.q[l] = qb.v[l];
    Big mq = mod(M, qb);
    if (prod)
        for (uint32_t l = 0; l < L; ++l)
            if ((*prod)[k * kMaxLimbs + l] != (l < mq.v.size() ? mq.v[l] : 0u))
                throw std::runtime_error("self test: device M mod q");
    Big neg(L);
    if (mq.bits() != 0) {
        neg = qb;
        sub_in(neg, mq);
    }
    for (uint32_t l = 0; l < L; ++l) qc.neg_m[l] = neg.v[l];
    Big mu = pow2_div((int)q_bits - 1 + 72, qb, 4);
    for (int i = 0; i < 4; ++i) qc.mu[i] = mu.v[i];
    if (qc.mu[3] != 0) throw std::runtime_error("Barrett constant wider than 96 bits");
    qc.q_bits = q_bits;
    qc.L = L;
}

// Device-chosen primes: every one checked here, and the product must meet
// the same coefficient bound the host rule stops at.
void adopt_primes(const std::vector<uint64_t>& primes, uint32_t N, uint32_t log_n, uint32_t q_bits, int bits,
                  size_t max_k, Big& M) {
    if (primes.empty() || primes.size() > max_k) throw std::runtime_error("self test: device prime count");
    M = Big::from_u64(1, 80);
    for (size_t j = 0; j < primes.size(); ++j) {
        const uint64_t p = primes[j];
        if (!is_prime(p) || (p - 1) % (2ull * N) != 0 || p >= (1ull << bits) || (j && p >= primes[j - 1]))
            throw std::runtime_error("self test: device prime search");
        M.mul_u64(p);
    }
    if ((uint32_t)M.bits() < log_n + 2 * q_bits + 4) throw std::runtime_error("self test: device primes too few");
}

void build62(HostTables& t, const Big& qb, const DeviceInit* init) {
    HostSet62& s = t.s62;
    Big M;
    if (init) {
        s.primes = init->primes62;
        adopt_primes(s.primes, t.N, t.log_n, t.q_bits, 62, (size_t)kMaxPrimes, M);
    } else {
        s.primes = pick_primes(t.N, t.log_n, t.q_bits, 62, (size_t)kMaxPrimes, M);
    }
    const uint32_t k = (uint32_t)s.primes.size(), N = t.N;
    if (t.host_twiddles) {
        s.tw_fwd.resize((size_t)k * N);
        s.tw_inv.resize((size_t)k * N);
    }
    s.tw_pow.resize((size_t)k * 2 * t.log_n);
    s.pc.resize(k);
    s.place.assign((size_t)k * kMaxLimbs, 0);
    s.cc.resize(k);
    // y_j = (M / p_j)^-1 mod p_j rides on the inverse transform's merged N^-1
    // constants, so the transform emits the CRT-weighted residue
    // v_j = (c mod p_j) y_j mod p_j that the CRT consumes.
    std::vector<uint64_t> y;
    crt_common(s.primes, M, qb, t.L, t.q_bits, s.miq, y, s.qc, init ? &init->prod62 : nullptr);
    for (uint32_t j = 0; j < k; ++j) {
        const uint64_t p = s.primes[j];
        RootTables r = roots(p, N, t.log_n, 64, t.host_twiddles, init ? init->psi62[j] : 0);
        r.nr = mulmod(r.nr, y[j], p);
        r.wr = mulmod(r.wr, y[j], p);
        for (uint32_t i = 0; i < t.log_n; ++i) {
            s.tw_pow[((size_t)j * 2 + 0) * t.log_n + i] = Tw{r.pf[i], shoup_w(r.pf[i], p, 64)};
            s.tw_pow[((size_t)j * 2 + 1) * t.log_n + i] = Tw{r.pi[i], shoup_w(r.pi[i], p, 64)};
        }
        for (uint32_t x = 0; x < N && t.host_twiddles; ++x) {
            s.tw_fwd[(size_t)j * N + x] = Tw{r.fwd[x], x ? shoup_w(r.fwd[x], p, 64) : 0};
            s.tw_inv[(size_t)j * N + x] = Tw{r.inv[x], x ? shoup_w(r.inv[x], p, 64) : 0};
        }
        PrimeConst& c = s.pc[j];
        c.p = p;
        c.p2 = 2 * p;
        c.mont_neg = neg_inv(p, 64);
        c.nr = r.nr;
        c.nr_s = shoup_w(r.nr, p, 64);
        c.wr = r.wr;
        c.wr_s = shoup_w(r.wr, p, 64);
        c.pad = 0;
        uint64_t pl = (uint64_t)((((u128)1) << 64) % p);  // 2^64 mod p: cancelled by REDC
        const uint64_t two32 = (1ull << 32) % p;
        for (uint32_t l = 0; l < t.L; ++l) {
            s.place[(size_t)j * kMaxLimbs + l] = pl;
            pl = mulmod(pl, two32, p);
        }
    }
    for (uint32_t j = 0; j < k; ++j) {
        const uint64_t p = s.primes[j];
        CrtConst& cc = s.cc[j];
        cc.y = y[j];
        cc.y_s = shoup_w(y[j], p, 64);
        const u128 mag = (~(u128)0) / p;  // floor(2^128 / p) (p odd, so 2^128 is no multiple)
        cc.mag_hi = (uint64_t)(mag >> 64);
        cc.mag_lo = (uint64_t)mag;
    }
}

// int8 MFMA CRT table.  With kappa = 4 j + u and raw bytes A[row][kappa]:
//   q-limb rows (lambda < LQ, byte b): byte 4 lambda + b of
//       m_ju = (2^(8u) ((M / p_j) mod q) R) mod q, R = 2^32, so that
//       sum_{lambda, b} 2^(32 lambda + 8 b) X[lambda, b] = sum_ju v_ju m_ju
//       = R sum_j v_j ((M / p_j) mod q)  (mod q),  v_ju byte u of v_j
//       (the kernel's Montgomery finish removes R);
//   fraction rows (lambda = LF + e): byte 4 e + b of 2^(8u) floor(2^56 / p_j),
//       so that the two pseudo-limbs hold F = sum_j v_j floor(2^56 / p_j)
//       exactly (F < 2^62);
//   ones rows (lambda = LF + 2, LF + 3, byte 0): stored value 1.
// The kernel contracts A - 128 with B - 128 (B = bytes of v, xor 0x80); with
// SA = row sums of A and SB = byte sum of the coefficient's v,
//   X = D + 128 SA + 128 SB - 128^2 Kdim,
// and D of a ones row is SB - 128 Kdim.  Per pseudo-limb the kernel forms
// Y = sum_b 2^(8b) D_b + 128 SB 0x01010101 + C_lambda = sum_b 2^(8b) X_b
// exactly, with C_lambda = sum_b 2^(8b) (128 SA - 128^2 Kdim): the limb sums
// then carry into the exact non-negative sum_t 2^(8t) X_t < 2^(32 L + 17).
void build_crt_mfma(HostTables& t, const Big& qb, const DeviceInit* init) {
    HostSet30& s = t.s30;
    const uint32_t L = t.L, k = (uint32_t)s.primes.size();
    // built for moduli that fill their L limbs to within 54 bits; narrower
    // ones (a short q padded to many limbs) keep the VALU CRT
    if (t.q_bits + 54 < 32 * L) return;
    const uint32_t rt = (L + 4 + 7) / 8, kc = (k + 7) / 8;
    const uint32_t lf = 8 * rt - 4, lq = std::min(lf, (uint32_t)kMaxLimbs);
    const uint32_t kbytes = 32 * kc;
    s.crt_rt = rt;
    s.crt_kc = kc;
    if (init && !init->crt_tab.empty()) {  // made on the device; the self tests emulate the contraction over it
        if (init->crt_tab.size() != (size_t)rt * 32 * kbytes || init->crt_c.size() != (size_t)8 * rt ||
            init->neg_m_r.size() != (size_t)kMaxLimbs)
            throw std::runtime_error("self test: device CRT table shape");
        s.crt_tab = init->crt_tab;
        s.crt_c = init->crt_c;
        std::memset(&s.crt_mc, 0, sizeof s.crt_mc);
        for (uint32_t l = 0; l < L; ++l) s.crt_mc.neg_m_r[l] = init->neg_m_r[l];
        s.crt_mc.q_inv = (uint32_t)neg_inv(qb.v[0], 32);
        return;
    }
    s.crt_tab.assign((size_t)rt * 32 * kbytes, 0);
    s.crt_c.assign((size_t)8 * rt, 0);
    auto row_of = [](uint32_t lam, uint32_t b) { return 32 * (lam / 8) + 8 * ((lam % 8) / 2) + 4 * (lam % 2) + b; };
    std::vector<int64_t> sa((size_t)rt * 32, 0);
    auto put = [&](uint32_t lam, uint32_t b, uint32_t kappa, uint32_t raw) {
        const uint32_t row = row_of(lam, b);
        s.crt_tab[(size_t)row * kbytes + kappa] = (int8_t)((int)raw - 128);
        sa[row] += raw;
    };
    for (uint32_t j = 0; j < 8 * kc; ++j) {
        Big m(L + 2);
        uint64_t g = 0;
        if (j < k) {
            for (uint32_t l = 0; l < L; ++l) m.v[l + 1] = s.miq[(size_t)j * kMaxLimbs + l];  // miq_j R
            m = mod(m, qb);
            m.v.resize(L + 1, 0u);
            g = (uint64_t)((((u128)1) << 56) / s.primes[j]);
        }
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t kappa = 4 * j + u;
            Big mr = j < k ? mod(m, qb) : Big(L);  // m_ju = 2^(8u) miq_j mod q
            for (uint32_t lam = 0; lam < lq; ++lam)
                for (uint32_t b = 0; b < 4; ++b) {
                    const uint32_t byte = lam < L ? (mr.v[lam] >> (8 * b)) & 0xffu : 0u;
                    put(lam, b, kappa, byte);
                }
            const uint64_t gu = g << (8 * u);  // < 2^51
            for (uint32_t e = 0; e < 2; ++e)
                for (uint32_t b = 0; b < 4; ++b) put(lf + e, b, kappa, (uint32_t)(gu >> (8 * (4 * e + b))) & 0xffu);
            for (uint32_t e = 2; e < 4; ++e) s.crt_tab[(size_t)row_of(lf + e, 0) * kbytes + kappa] = 1;
            if (j < k) m.mul_u64(256);
        }
    }
    const int64_t off = 128ll * 128ll * kbytes;  // 128^2 Kdim
    for (uint32_t lam = 0; lam < 8 * rt; ++lam) {  // exact: sum_b 2^(8b) (128 SA - 128^2 Kdim), |.| < 2^48
        if (lam >= lq && lam != lf && lam != lf + 1) continue;
        int64_t c = 0;
        for (uint32_t b = 0; b < 4; ++b) c += (128 * sa[row_of(lam, b)] - off) * ((int64_t)1 << (8 * b));
        s.crt_c[lam] = c;
    }
    // Montgomery finish constants: ((-M) mod q) R mod q and -q^-1 mod 2^32
    Big nm(L + 1);
    for (uint32_t l = 0; l < L; ++l) nm.v[l + 1] = s.qc.neg_m[l];
    nm = mod(nm, qb);
    std::memset(&s.crt_mc, 0, sizeof s.crt_mc);
    for (uint32_t l = 0; l < L; ++l) s.crt_mc.neg_m_r[l] = nm.v[l];
    s.crt_mc.q_inv = (uint32_t)neg_inv(qb.v[0], 32);
}

// Rows k .. 3k - 1 of the 30-bit prime constants: variant v (1, 2) of row j
// is row j with the final inverse stage's factors nr, wr times 2^v, for inner
// blocks that merge v levels of the transform into their pointwise product
// and so leave their output scaled by 2^-v (merge30 in nc_kernels.hip).
void add_merge_variants(std::vector<PrimeConst32>& pc, size_t k) {
    pc.resize((size_t)kMergeVariants30 * k);
    for (size_t v = 1; v < (size_t)kMergeVariants30; ++v)
        for (size_t j = 0; j < k; ++j) {
            PrimeConst32 c = pc[j];
            const uint64_t p = c.p, f = 1ull << v;
            c.nr = (uint32_t)mulmod(c.nr, f, p);
            c.nr_s = (uint32_t)shoup_w(c.nr, p, 32);
            c.wr = (uint32_t)mulmod(c.wr, f, p);
            c.wr_s = (uint32_t)shoup_w(c.wr, p, 32);
            pc[v * k + j] = c;
        }
}

void build30(HostTables& t, const Big& qb, const DeviceInit* init) {
    HostSet30& s = t.s30;
    Big M;
    if (init) {
        s.primes = init->primes30;
        adopt_primes(s.primes, t.N, t.log_n, t.q_bits, 30, (size_t)kMaxPrimes30, M);
    } else {
        s.primes = pick_primes(t.N, t.log_n, t.q_bits, 30, (size_t)kMaxPrimes30, M);
    }
    const uint32_t k = (uint32_t)s.primes.size(), N = t.N;
    if (t.host_twiddles) {
        s.tw_fwd.resize((size_t)k * N);
        s.tw_inv.resize((size_t)k * N);
    }
    s.tw_pow.resize((size_t)k * 2 * t.log_n);
    s.pc.resize(k);
    s.place.assign((size_t)k * kMaxLimbs, 0);
    s.cc.resize(k);
    std::vector<uint64_t> y;  // CRT weight, merged as in build62
    crt_common(s.primes, M, qb, t.L, t.q_bits, s.miq, y, s.qc, init ? &init->prod30 : nullptr);
    for (uint32_t j = 0; j < k; ++j) {
        const uint64_t p = s.primes[j];
        RootTables r = roots(p, N, t.log_n, 32, t.host_twiddles, init ? init->psi30[j] : 0);
        r.nr = mulmod(r.nr, y[j], p);
        r.wr = mulmod(r.wr, y[j], p);
        for (uint32_t i = 0; i < t.log_n; ++i) {
            s.tw_pow[((size_t)j * 2 + 0) * t.log_n + i] = Tw32{(uint32_t)r.pf[i], (uint32_t)shoup_w(r.pf[i], p, 32)};
            s.tw_pow[((size_t)j * 2 + 1) * t.log_n + i] = Tw32{(uint32_t)r.pi[i], (uint32_t)shoup_w(r.pi[i], p, 32)};
        }
        for (uint32_t x = 0; x < N && t.host_twiddles; ++x) {
            // forward twiddles stored negated (2^32 - w) for the one-multiply-add butterfly
            s.tw_fwd[(size_t)j * N + x] = Tw32{0u - (uint32_t)r.fwd[x], x ? (uint32_t)shoup_w(r.fwd[x], p, 32) : 0u};
            s.tw_inv[(size_t)j * N + x] = Tw32{(uint32_t)r.inv[x], x ? (uint32_t)shoup_w(r.inv[x], p, 32) : 0u};
        }
        PrimeConst32& c = s.pc[j];
        c.p = (uint32_t)p;
        c.p2 = (uint32_t)(2 * p);
        c.mont_neg = (uint32_t)neg_inv(p, 32);
        c.nr = (uint32_t)r.nr;
        c.nr_s = (uint32_t)shoup_w(r.nr, p, 32);
        c.wr = (uint32_t)r.wr;
        c.wr_s = (uint32_t)shoup_w(r.wr, p, 32);
        c.c32 = (uint32_t)((1ull << 32) % p);
        uint64_t pl = (1ull << 32) % p;  // 2^32 mod p: cancelled by the final REDC
        for (uint32_t l = 0; l < t.L; ++l) {
            s.place[(size_t)j * kMaxLimbs + l] = (uint32_t)pl;
            pl = mulmod(pl, c.c32, p);
        }
    }
    add_merge_variants(s.pc, k);
    // int8 MFMA decomposition table: T_m = 2^(8m) 2^32 mod p_j written in
    // balanced base-256 digits t_um in [-128, 127] (sum_u 2^(8u) t_um = T_m),
    // which are valid signed int8 as they are.  The kernel contracts them with
    // the coefficient bytes offset to signed (a_m - 128):
    //   D_u = sum_m (a_m - 128) t_um,  X_u = sum_m a_m t_um = D_u + 128 sum_m t_um,
    // so V_j = sum_m a_m T_m = sum_u 2^(8u) X_u.  The row bias 128 sum_m t_um
    // enters the contraction as its accumulator input, so the kernel reads
    // X_u directly.
    if (t.L % 4 == 0) {
        const uint32_t tiles = (k + 7) / 8, nbytes = 4 * t.L;
        s.mfma_tab.assign((size_t)tiles * 32 * kMfmaK, 0);
        s.mfma_bias.assign((size_t)tiles * 32, 0);
        for (uint32_t j = 0; j < k; ++j) {
            const uint64_t p = s.primes[j];
            const size_t row0 = (size_t)(j / 8) * 32 + 4 * (j % 8);
            uint64_t place = (1ull << 32) % p;  // 2^(8m) 2^32 mod p, m = 0
            int64_t bias[4] = {0, 0, 0, 0};
            for (uint32_t m = 0; m < nbytes; ++m) {
                int64_t rest = (int64_t)place;
                for (int u = 0; u < 4; ++u) {
                    int64_t d = rest & 0xff;
                    if (d >= 128) d -= 256;
                    rest = (rest - d) >> 8;
                    s.mfma_tab[(row0 + u) * kMfmaK + m] = (int8_t)d;
                    bias[u] += 128 * d;
                }
                if (rest != 0) throw std::runtime_error("MFMA table digit overflow");
                place = mulmod(place, 256, p);
            }
            for (int u = 0; u < 4; ++u) s.mfma_bias[row0 + u] = (int32_t)bias[u];
        }
    }
    build_crt_mfma(t, qb, init);
    for (uint32_t j = 0; j < k; ++j) {
        const uint64_t p = s.primes[j];
        CrtConst32& cc = s.cc[j];
        cc.y = (uint32_t)y[j];
        cc.y_s = (uint32_t)shoup_w(y[j], p, 32);
        const u128 g = (((u128)1) << 96) / p;
        cc.g_hi = (uint32_t)(g >> 64);
        cc.pad = 0;
        cc.g_lo = (uint64_t)g;
    }
}

}  // namespace

bool select_primes(const std::vector<std::pair<uint64_t, uint64_t>>& cands, uint32_t log_n, uint32_t q_bits,
                   size_t max_k, std::vector<uint64_t>& primes, std::vector<uint64_t>& psi) {
    const uint32_t bound_bits = log_n + 2 * q_bits + 4;
    Big M = Big::from_u64(1, 80);
    primes.clear();
    psi.clear();
    for (const auto& c : cands) {
        if ((uint32_t)M.bits() >= bound_bits) break;
        if (!c.second) continue;
        if (primes.size() >= max_k) return false;
        primes.push_back(c.first);
        psi.push_back(c.second);
        M.mul_u64(c.first);
    }
    return (uint32_t)M.bits() >= bound_bits;
}

void validate_modulus(uint32_t L, const uint32_t* q) {
    if (L == 0 || L > (uint32_t)kMaxLimbs) throw std::invalid_argument("L must be in [1, 32]");
    if (!q || !(q[0] & 1u)) throw std::invalid_argument("q must be an odd modulus of L limbs");
    Big qb(L);
    for (uint32_t l = 0; l < L; ++l) qb.v[l] = q[l];
    const int bits = qb.bits();
    if (bits < 2) throw std::invalid_argument("q must be an odd modulus > 1");
    // fast-CRT width (solve.cu:1339-1341): bits(q) + 70 <= 34 * 32
    if (bits + 70 > 1088) throw std::invalid_argument("q wider than 1018 bits is not supported");
}

HostTables build_tables(uint32_t N, uint32_t L, const uint32_t* q, bool host_twiddles, const DeviceInit* init) {
    // validation (solve.cu:1167-1172; N cap raised to 65536)
    if (N < 2 || N > (1u << kMaxLogN) || (N & (N - 1))) throw std::invalid_argument("N must be a power of two from 2 to 65536");
    validate_modulus(L, q);
    HostTables t;
    t.N = N;
    t.L = L;
    t.host_twiddles = host_twiddles;
    while ((1u << t.log_n) < N) ++t.log_n;
    Big qb(L);
    for (uint32_t l = 0; l < L; ++l) qb.v[l] = q[l];
    t.q_bits = (uint32_t)qb.bits();
    if (t.q_bits < 2) throw std::invalid_argument("q must be an odd modulus > 1");
    // fast-CRT width (solve.cu:1339-1341): bits(q) + 70 <= 34 * 32
    if (t.q_bits + 70 > 1088) throw std::invalid_argument("q wider than 1018 bits is not supported");
    if (init && (init->psi62.size() != init->primes62.size() || init->psi30.size() != init->primes30.size()))
        throw std::runtime_error("self test: device psi count");
    build62(t, qb, init);
    build30(t, qb, init);
    return t;
}

// Degree-N tables on a shared set (the master's tables, built for its
// largest degree Nm with primes p = 1 mod 2 Nm): psi_N = psi_Nm^(Nm / N), so
// psi_N^(2^i) is the master's psi^(2^(i + log Nm - log N)) -- the tail of its
// powers table -- and degree N's twiddle entry x < N is the master's entry x
// (bitrev over log Nm bits of x < N is bitrev over log N bits shifted up by
// log Nm - log N).  Only the per-prime constants that carry N^-1 and
// psi_N^(-N/2) are recomputed; every other table is N-independent.
HostTables derive_tables(const HostTables& m, uint32_t N) {
    if (N < 2 || N > m.N || (N & (N - 1))) throw std::invalid_argument("N must be a power of two from 2 to the set's largest degree");
    if (m.host_twiddles) throw std::invalid_argument("a shared set keeps its twiddles on the device");
    HostTables t = m;
    t.N = N;
    t.log_n = 0;
    while ((1u << t.log_n) < N) ++t.log_n;
    const uint32_t sh = m.log_n - t.log_n;
    auto tail = [&](const auto& src, auto& dst, size_t k) {
        dst.resize(k * 2 * t.log_n);
        for (size_t j = 0; j < k; ++j)
            for (size_t dir = 0; dir < 2; ++dir)
                for (uint32_t i = 0; i < t.log_n; ++i)
                    dst[(j * 2 + dir) * t.log_n + i] = src[(j * 2 + dir) * m.log_n + i + sh];
    };
    tail(m.s62.tw_pow, t.s62.tw_pow, m.s62.primes.size());
    tail(m.s30.tw_pow, t.s30.tw_pow, m.s30.primes.size());
    for (size_t j = 0; j < m.s62.primes.size(); ++j) {
        const uint64_t p = m.s62.primes[j];
        const uint64_t R = (uint64_t)((((u128)1) << 64) % p);
        const uint64_t nr = mulmod(mulmod(invmod(N % p, p), R, p), m.s62.cc[j].y, p);
        const uint64_t wr = mulmod(t.s62.tw_pow[(j * 2 + 1) * t.log_n + t.log_n - 1].w, nr, p);  // psi_N^-(N/2)
        PrimeConst& c = t.s62.pc[j];
        c.nr = nr;
        c.nr_s = shoup_w(nr, p, 64);
        c.wr = wr;
        c.wr_s = shoup_w(wr, p, 64);
    }
    for (size_t j = 0; j < m.s30.primes.size(); ++j) {
        const uint64_t p = m.s30.primes[j];
        const uint64_t R = (1ull << 32) % p;
        const uint64_t nr = mulmod(mulmod(invmod(N % p, p), R, p), m.s30.cc[j].y, p);
        const uint64_t wr = mulmod(t.s30.tw_pow[(j * 2 + 1) * t.log_n + t.log_n - 1].w, nr, p);
        PrimeConst32& c = t.s30.pc[j];
        c.nr = (uint32_t)nr;
        c.nr_s = (uint32_t)shoup_w(nr, p, 32);
        c.wr = (uint32_t)wr;
        c.wr_s = (uint32_t)shoup_w(wr, p, 32);
    }
    add_merge_variants(t.s30.pc, m.s30.primes.size());
    return t;
}

Tw twiddle_entry62(const HostTables& t, uint32_t j, uint32_t x, bool inverse) {
    if (x == 0) return Tw{0, 0};
    const uint64_t p = t.s62.primes[j];
    std::vector<uint64_t> sq(t.log_n);
    for (uint32_t i = 0; i < t.log_n; ++i) sq[i] = t.s62.tw_pow[((size_t)j * 2 + inverse) * t.log_n + i].w;
    const uint64_t w = pow_from_table(sq, bitrev(x, t.log_n), p);
    return Tw{w, shoup_w(w, p, 64)};
}

Tw32 twiddle_entry30(const HostTables& t, uint32_t j, uint32_t x, bool inverse) {
    if (x == 0) return Tw32{0u, 0u};
    const uint64_t p = t.s30.primes[j];
    std::vector<uint64_t> sq(t.log_n);
    for (uint32_t i = 0; i < t.log_n; ++i) sq[i] = t.s30.tw_pow[((size_t)j * 2 + inverse) * t.log_n + i].w;
    const uint64_t w = pow_from_table(sq, bitrev(x, t.log_n), p);
    return Tw32{inverse ? (uint32_t)w : 0u - (uint32_t)w, (uint32_t)shoup_w(w, p, 32)};
}

namespace {

// Host emulation of the int8 MFMA decomposition (decompose30_mfma_kernel):
// for extreme and random coefficients, X_u = D_u + bias, the kernel's int32
// byte pairs lo = X_0 + 2^8 X_1, hi = X_2 + 2^8 X_3 (which must not wrap),
// V = lo + 2^16 hi and its REDC must equal the coefficient mod p_j.
void check_decompose_mfma(const HostTables& t, std::mt19937_64& rng) {
    const HostSet30& s = t.s30;
    const uint32_t L = t.L, nbytes = 4 * L;
    std::vector<uint32_t> limbs(L);
    for (int trial = 0; trial < 4; ++trial) {
        for (uint32_t l = 0; l < L; ++l) limbs[l] = trial == 0 ? 0xffffffffu : trial == 1 ? 0u : (uint32_t)rng();
        for (uint32_t j = 0; j < s.primes.size(); ++j) {
            const uint64_t p = s.primes[j];
            const size_t row0 = (size_t)(j / 8) * 32 + 4 * (j % 8);
            int64_t x[4];
            for (int u = 0; u < 4; ++u) {
                int64_t d = s.mfma_bias[row0 + u];
                for (uint32_t m = 0; m < (uint32_t)kMfmaK; ++m) {
                    const int a = m < nbytes ? (int)((limbs[m / 4] >> (8 * (m % 4))) & 0xffu) : 0;
                    d += (int64_t)(a - 128) * s.mfma_tab[(row0 + u) * kMfmaK + m];
                }
                x[u] = d;
            }
            const int64_t lo = x[0] + 256 * x[1], hi = x[2] + 256 * x[3];
            if (lo != (int32_t)lo || hi != (int32_t)hi) throw std::runtime_error("self test: MFMA decomposition range");
            const int64_t v = lo + 65536 * hi;
            const uint64_t V = (uint64_t)v;
            const uint32_t mneg = s.pc[j].mont_neg;
            const uint32_t m32 = (uint32_t)V * mneg;
            uint64_t r = (uint64_t)(((u128)m32 * p + V) >> 32);
            if (r >= p) r -= p;
            uint64_t horner = 0;
            for (int l = (int)L - 1; l >= 0; --l) horner = (uint64_t)((((u128)horner) << 32 | limbs[l]) % p);
            if (v < 0 || r != horner) throw std::runtime_error("self test: MFMA decomposition table");
        }
    }
}

// Host emulation of the int8 MFMA CRT contraction (crt30_mfma_kernel) on
// extreme and random CRT-weighted residues, canonical ([0, p_j)) and, for
// the shape the fused inverse stages run at (L = 28, rt = 4, kc = 8, so
// k <= 64), lazy ([0, 2 p_j), what those stages hand over): the q limbs must
// agree with sum_j v_j ((M / p_j) mod q) mod q, empty limbs must be exactly
// zero and the fraction limbs must equal sum_j v_j floor(2^56 / p_j), below
// 2^63 (so the kernel's 64-bit fraction and its K = round(F / 2^56) < 128
// hold).
void check_crt_mfma(const HostTables& t, std::mt19937_64& rng) {
    const HostSet30& s = t.s30;
    const uint32_t L = t.L, k = (uint32_t)s.primes.size();
    const uint32_t rt = s.crt_rt, kbytes = 32 * s.crt_kc;
    const uint32_t lf = 8 * rt - 4, lq = std::min(lf, (uint32_t)kMaxLimbs);
    Big qb(L);
    for (uint32_t l = 0; l < L; ++l) qb.v[l] = s.qc.q[l];
    auto row_of = [](uint32_t lam, uint32_t b) { return 32 * (lam / 8) + 8 * ((lam % 8) / 2) + 4 * (lam % 2) + b; };
    const bool fused_shape = L == 28 && rt == 4 && s.crt_kc == 8;
    for (int trial = 0; trial < (fused_shape ? 6 : 4); ++trial) {
        std::vector<uint32_t> v(8 * s.crt_kc, 0u);
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t p = (uint32_t)s.primes[j];
            v[j] = trial == 0 ? p - 1 : trial == 1 ? 0u : trial == 4 ? 2 * p - 1 : trial == 5 ? (uint32_t)(rng() % (2 * p))
                                                                                    : (uint32_t)(rng() % p);
        }
        std::vector<int> bq(kbytes);
        int64_t sb = 0;
        for (uint32_t kap = 0; kap < kbytes; ++kap) {
            const uint32_t byte = (v[kap / 4] >> (8 * (kap % 4))) & 0xffu;
            bq[kap] = (int)byte - 128;
            sb += byte;
        }
        auto dot = [&](uint32_t row) {
            int64_t d = 0;
            for (uint32_t kap = 0; kap < kbytes; ++kap) d += (int64_t)s.crt_tab[(size_t)row * kbytes + kap] * bq[kap];
            return d;
        };
        if (dot(row_of(lf + 2, 0)) != sb - 128 * (int64_t)kbytes || dot(row_of(lf + 3, 0)) != sb - 128 * (int64_t)kbytes)
            throw std::runtime_error("self test: CRT table ones rows");
        auto Y = [&](uint32_t lam) {
            int64_t y = 0;
            for (uint32_t b = 0; b < 4; ++b) y += dot(row_of(lam, b)) * ((int64_t)1 << (8 * b));
            return y + 128 * sb * 0x01010101ll + s.crt_c[lam];
        };
        std::vector<uint32_t> acc(lq + 2, 0u);
        int64_t carry = 0;
        for (uint32_t lam = 0; lam < lq; ++lam) {
            const int64_t y = Y(lam);
            if (lam >= L && y != 0) throw std::runtime_error("self test: CRT table empty limb");
            const int64_t x = y + carry;
            acc[lam] = (uint32_t)x;
            carry = x >> 32;
        }
        if (carry < 0) throw std::runtime_error("self test: CRT table sign");
        acc[lq] = (uint32_t)carry;
        acc[lq + 1] = (uint32_t)(carry >> 32);
        const int64_t f0 = Y(lf), f1 = Y(lf + 1);
        u128 fwant = 0;
        for (uint32_t j = 0; j < k; ++j) fwant += (u128)v[j] * (uint64_t)((((u128)1) << 56) / s.primes[j]);
        if (f0 < 0 || f1 < 0 || (u128)(uint64_t)f0 + ((u128)(uint64_t)f1 << 32) != fwant || (fwant >> 63) != 0)
            throw std::runtime_error("self test: CRT table fraction");
        Big want(L + 3);
        for (uint32_t j = 0; j < k; ++j) {
            Big term(L + 3);
            for (uint32_t l = 0; l < L; ++l) term.v[l] = s.miq[(size_t)j * kMaxLimbs + l];
            term.mul_u64(v[j]);
            uint64_t c = 0;
            for (uint32_t l = 0; l < L + 3; ++l) {
                const uint64_t x = (uint64_t)want.v[l] + term.v[l] + c;
                want.v[l] = (uint32_t)x;
                c = x >> 32;
            }
        }
        Big wr(L + 4);  // want R
        for (uint32_t l = 0; l < L + 3; ++l) wr.v[l + 1] = want.v[l];
        Big got(lq + 2);
        got.v = acc;
        if (cmp(mod(got, qb), mod(wr, qb)) != 0) throw std::runtime_error("self test: CRT table sum");
        // the kernel adds K ((-M) mod q) R mod q (K < 128) and runs one 2^32
        // Montgomery step, whose result is below 2q only for sums below
        // (2^32 - 128) q: the byte-weighted sum stays below 2^16 q
        if (got.bits() > (int)t.q_bits + 16) throw std::runtime_error("self test: CRT table range");
        // the kernel's finish on these limbs: K = round(F / 2^56), x = sum +
        // K neg_m_r, one 2^32 Montgomery step, one conditional subtraction;
        // the result must be c mod q, c = sum_j v_j (M / p_j) - K M, i.e.
        // (sum_j v_j ((M / p_j) mod q) + K ((-M) mod q)) mod q
        const uint32_t K = (uint32_t)(fwant >> 56) + (uint32_t)((fwant >> 55) & 1u);
        std::vector<uint32_t> x = acc;
        uint64_t cy = 0;
        for (uint32_t l = 0; l < lq + 2; ++l) {
            const uint64_t z = (uint64_t)x[l] + (l < L ? (uint64_t)K * s.crt_mc.neg_m_r[l] : 0u) + cy;
            x[l] = (uint32_t)z;
            cy = z >> 32;
        }
        const uint32_t mq = x[0] * s.crt_mc.q_inv;
        for (uint32_t l = 0; l < lq + 2; ++l) {
            const uint64_t z = (uint64_t)x[l] + (l < L ? (uint64_t)mq * s.qc.q[l] : 0u) + cy;
            x[l] = (uint32_t)z;
            cy = z >> 32;
        }
        if (cy || x[0]) throw std::runtime_error("self test: CRT finish step");
        Big r(lq + 1), q2(lq + 1);
        for (uint32_t l = 0; l <= lq; ++l) r.v[l] = x[l + 1];
        for (uint32_t l = 0; l < L; ++l) q2.v[l] = s.qc.q[l];
        q2.mul_u64(2);
        if (cmp(r, q2) >= 0) throw std::runtime_error("self test: CRT finish range");
        if (cmp(r, qb) >= 0) sub_in(r, qb);
        Big kn(L + 3);
        for (uint32_t l = 0; l < L; ++l) kn.v[l] = s.qc.neg_m[l];
        kn.mul_u64(K);
        uint64_t c2 = 0;
        for (uint32_t l = 0; l < L + 3; ++l) {
            const uint64_t z = (uint64_t)want.v[l] + kn.v[l] + c2;
            want.v[l] = (uint32_t)z;
            c2 = z >> 32;
        }
        if (c2 || cmp(mod(want, qb), mod(r, qb)) != 0 || r.bits() > (int)t.q_bits)
            throw std::runtime_error("self test: CRT finish value");
    }
}

}  // namespace

// Known-answer checks of both prime sets, run at plan creation (the
// analogue of solve.cu:1151-1166, 1285-1314, 1370-1415 for this formulation).
void self_test(const HostTables& t) {
    std::mt19937_64 rng(0x5eed);
    auto check_prime = [&](uint64_t p, uint64_t tw1, uint64_t tw1_inv, int bits) {
        if (!is_prime(p) || (p - 1) % (2ull * t.N) != 0 || p >= (1ull << bits))
            throw std::runtime_error("self test: prime rule violated");
        if (t.N > 1) {  // tw[1] = psi^(N/2) squares to -1
            if (mulmod(tw1, tw1, p) != p - 1) throw std::runtime_error("self test: psi order");
            if (mulmod(tw1, tw1_inv, p) != 1) throw std::runtime_error("self test: psi inverse");
        }
    };
    uint32_t limbs[kMaxLimbs];
    for (uint32_t j = 0; j < t.s62.primes.size(); ++j) {
        const PrimeConst& c = t.s62.pc[j];
        const uint64_t p = c.p;
        // psi^(N/2) = psi^bitrev(1), the first stage's twiddle
        const size_t top = t.log_n - 1;
        check_prime(p, t.s62.tw_pow[((size_t)j * 2 + 0) * t.log_n + top].w,
                    t.s62.tw_pow[((size_t)j * 2 + 1) * t.log_n + top].w, 62);
        if (t.host_twiddles && t.N > 1 &&
            (t.s62.tw_fwd[(size_t)j * t.N + 1].w != t.s62.tw_pow[((size_t)j * 2 + 0) * t.log_n + top].w ||
             t.s62.tw_inv[(size_t)j * t.N + 1].w != t.s62.tw_pow[((size_t)j * 2 + 1) * t.log_n + top].w))
            throw std::runtime_error("self test: twiddle table");
        if ((uint64_t)(p * (0 - c.mont_neg)) != 1) throw std::runtime_error("self test: Montgomery constant");
        for (int trial = 0; trial < 16; ++trial) {  // Shoup lazy product in [0, 2p), exact mod p
            const uint64_t a = rng(), w = rng() % p, ws = shoup_w(w, p, 64);
            const uint64_t lazy = a * w - (uint64_t)(((u128)a * ws) >> 64) * p;
            if (lazy >= 2 * p || lazy % p != mulmod(a % p, w, p)) throw std::runtime_error("self test: Shoup product");
        }
        for (uint32_t l = 0; l < t.L; ++l) limbs[l] = (uint32_t)rng();
        u128 s = 0;  // place table: REDC(sum limb 2^(32l) 2^64) equals Horner
        for (uint32_t l = 0; l < t.L; ++l) s += (u128)limbs[l] * t.s62.place[(size_t)j * kMaxLimbs + l];
        const uint64_t lo = (uint64_t)s;
        uint64_t redc = (uint64_t)(s >> 64) + (uint64_t)(((u128)(lo * c.mont_neg) * p) >> 64) + (lo != 0);
        if (redc >= p) redc -= p;
        uint64_t horner = 0;
        for (int l = (int)t.L - 1; l >= 0; --l) horner = (uint64_t)((((u128)horner) << 32 | limbs[l]) % p);
        if (redc != horner) throw std::runtime_error("self test: digit sum does not match the modulus");
    }
    for (uint32_t j = 0; j < t.s30.primes.size(); ++j) {
        const PrimeConst32& c = t.s30.pc[j];
        const uint32_t p = c.p;
        const size_t top = t.log_n - 1;
        check_prime(p, t.s30.tw_pow[((size_t)j * 2 + 0) * t.log_n + top].w,
                    t.s30.tw_pow[((size_t)j * 2 + 1) * t.log_n + top].w, 30);
        if (t.host_twiddles && t.N > 1 &&
            (0u - t.s30.tw_fwd[(size_t)j * t.N + 1].w != t.s30.tw_pow[((size_t)j * 2 + 0) * t.log_n + top].w ||
             t.s30.tw_inv[(size_t)j * t.N + 1].w != t.s30.tw_pow[((size_t)j * 2 + 1) * t.log_n + top].w))
            throw std::runtime_error("self test: twiddle table");
        if ((uint32_t)(p * (0u - c.mont_neg)) != 1u) throw std::runtime_error("self test: Montgomery constant");
        // the merge variants (rows v k + j): the same prime, the final stage's
        // factors times 2^v, each Shoup quotient matching its factor
        const size_t k30 = t.s30.primes.size();
        if (t.s30.pc.size() != (size_t)kMergeVariants30 * k30) throw std::runtime_error("self test: merge variants");
        for (int v = 1; v < kMergeVariants30; ++v) {
            const PrimeConst32& cv = t.s30.pc[(size_t)v * k30 + j];
            if (cv.p != p || cv.p2 != c.p2 || cv.mont_neg != c.mont_neg || cv.c32 != c.c32 ||
                mulmod(cv.nr, p - (p - 1) / 2, p) != mulmod(c.nr, 1ull << (v - 1), p) ||
                mulmod(cv.wr, p - (p - 1) / 2, p) != mulmod(c.wr, 1ull << (v - 1), p) ||
                cv.nr_s != (uint32_t)shoup_w(cv.nr, p, 32) || cv.wr_s != (uint32_t)shoup_w(cv.wr, p, 32))
                throw std::runtime_error("self test: merge variants");
        }
        for (int trial = 0; trial < 16; ++trial) {
            const uint32_t a = (uint32_t)rng(), w = (uint32_t)(rng() % p), ws = (uint32_t)shoup_w(w, p, 32);
            const uint32_t lazy = a * w - (uint32_t)(((uint64_t)a * ws) >> 32) * p;
            if (lazy >= 2 * p || lazy % p != mulmod(a % p, w, p)) throw std::runtime_error("self test: Shoup product");
        }
        // decomposition fold + REDC equals Horner
        for (uint32_t l = 0; l < t.L; ++l) limbs[l] = (uint32_t)rng();
        uint64_t s = 0;
        for (uint32_t l = 0; l < t.L; l += 2) {
            s += (uint64_t)limbs[l] * t.s30.place[(size_t)j * kMaxLimbs + l];
            if (l + 1 < t.L) s += (uint64_t)limbs[l + 1] * t.s30.place[(size_t)j * kMaxLimbs + l + 1];
            s = (s >> 32) * c.c32 + (uint32_t)s;
        }
        s = (s >> 32) * c.c32 + (uint32_t)s;
        const uint32_t lo = (uint32_t)s;
        uint32_t redc = (uint32_t)(s >> 32) + (uint32_t)(((uint64_t)(lo * c.mont_neg) * p) >> 32) + (lo != 0);
        if (redc >= p) redc -= p;
        uint64_t horner = 0;
        for (int l = (int)t.L - 1; l >= 0; --l) horner = (uint64_t)((((u128)horner) << 32 | limbs[l]) % p);
        if (redc != horner) throw std::runtime_error("self test: digit sum does not match the modulus");
    }
    if (!t.s30.mfma_tab.empty()) check_decompose_mfma(t, rng);
    if (!t.s30.crt_tab.empty()) check_crt_mfma(t, rng);
}

}  // namespace nc
