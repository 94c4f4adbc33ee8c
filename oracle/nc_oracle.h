/*
 * CPU oracle for the negacyclic product c = a*b mod (X^N + 1, q).
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker (or, for cpu_baseline, as the timed
 * CPU restatement of the reference algorithm).
 *
 * Two independent formulations live here:
 *   1. orc_schoolbook*: the exact big-integer schoolbook negacyclic product
 *      (the mathematical definition; no primes, no transforms).
 *   2. orc_plan_* / orc_multiply: a restatement, in plain C, of the
 *      reference's RNS path (solve.cu): the prime rule, the Montgomery /
 *      Shoup / Barrett word arithmetic, the psi-twisted DIT NTT, the
 *      pointwise product, the untwist and the fast CRT with its fixed-point
 *      quotient and centred lift; plus the (dead in the reference) Garner CRT
 *      as a cross-check.  Every function cites the solve.cu lines it follows.
 *
 * Layouts follow the reference boundary (fherma.h / negacyclic.fhk):
 *   polynomials: u32[N][L], limb l of coefficient i at i*L + l, little endian.
 *   residues:    u64[k][N], channel (prime) major, canonical in [0, p_j).
 */
#ifndef NC_ORACLE_H
#define NC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define ORC_OK 0
#define ORC_EINVAL -1
#define ORC_ENOMEM -2
#define ORC_ESELFTEST -3
#define ORC_ENOPRIMES -4

/* ---- exact schoolbook (definition) ------------------------------------ */
/* c = a*b mod (X^N+1, q), exact. threads<=0 means all OpenMP threads. */
int orc_schoolbook(uint32_t N, uint32_t L, const uint32_t* q, const uint32_t* a, const uint32_t* b,
                   uint32_t* c, int threads);
/* Only the coefficients listed in idx[0..count). out is count x L. */
int orc_schoolbook_coeffs(uint32_t N, uint32_t L, const uint32_t* q, const uint32_t* a, const uint32_t* b,
                          const uint32_t* idx, uint32_t count, uint32_t* out, int threads);

/* ---- seeded synthetic inputs ----------------------------------------- */
/* SplitMix64 stream from `seed`; each candidate coefficient draws
 * ceil(L/2) u64 words (word w -> limbs 2w, 2w+1), the top limb is masked to
 * bits(q) - 32(L-1) bits, and candidates >= q are rejected (redrawn). */
void orc_gen_poly(uint64_t seed, uint32_t N, uint32_t L, const uint32_t* q, uint32_t* out);

/* ---- restated reference path ------------------------------------------ */
typedef struct orc_plan orc_plan;

int orc_plan_create(uint32_t N, uint32_t L, const uint32_t* q, orc_plan** out);
void orc_plan_destroy(orc_plan* plan);
uint32_t orc_plan_channels(const orc_plan* plan);
void orc_plan_primes(const orc_plan* plan, uint64_t* out);
/* residues r[j][i] = a_i mod p_j (canonical, no Montgomery factor, no twist) */
int orc_residues(const orc_plan* plan, const uint32_t* poly, uint64_t* out);
/* product residues: c_int[i] mod p_j, where c_int is the exact integer
 * negacyclic product of a and b (before reduction mod q). */
int orc_product_residues(const orc_plan* plan, const uint32_t* a, const uint32_t* b, uint64_t* out,
                         int threads);
/* fast CRT (solve.cu:752-836) from canonical residues to L limbs mod q */
int orc_crt_fast(const orc_plan* plan, const uint64_t* residues, uint32_t* c, int threads);
/* Garner mixed-radix CRT (solve.cu:838-906) from canonical residues */
int orc_crt_garner(const orc_plan* plan, const uint64_t* residues, uint32_t* c, int threads);
/* whole run-equivalent (solve.cu:1489-1528): a, b -> c */
int orc_multiply(const orc_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, int threads);
/* per-coefficient CRT decisions of the fast CRT: kappa estimate sum_hi and
 * the centred-lift bit (sum_lo >= 2^63), for stage-level checks. */
int orc_crt_decisions(const orc_plan* plan, const uint64_t* residues, uint32_t* sum_hi, uint8_t* negative);

/* ---- known-answer self tests of the reference's init (solve.cu:1151-1166,
 * 1285-1314, 1370-1415); returns ORC_OK or the number of the failed check. */
int orc_selftest(const orc_plan* plan);

/* ---- word primitives, exported for unit tests ------------------------- */
uint64_t orc_mul_mod(uint64_t a, uint64_t b, uint64_t p);
uint64_t orc_mont_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t nprime);
uint64_t orc_shoup_mul(uint64_t a, uint64_t w, uint64_t w_shoup, uint64_t p);
uint64_t orc_reduce_128(uint64_t hi, uint64_t lo, uint64_t p, uint64_t magic_hi, uint64_t magic_lo);
uint64_t orc_barrett_reduce(uint64_t hi, uint64_t lo, uint64_t p, uint64_t mu);
uint64_t orc_barrett_mu(uint64_t p);
uint64_t orc_inv_mod_2_64(uint64_t odd);
uint64_t orc_pow_mod(uint64_t base, uint64_t exp, uint64_t p);
int orc_is_prime_u64(uint64_t n);
uint64_t orc_find_psi(uint64_t p, uint32_t n);
uint64_t orc_div_pow2_64(uint64_t p, uint64_t* remainder);
/* 34-limb x mod q by the reference's Barrett (q of 868 bits, L = 28) */
void orc_mod_q_barrett(uint32_t* acc34, const uint32_t* q28, const uint32_t* mu8);
/* generic shift-subtract x mod q over `wide` limbs */
void orc_mod_q(uint32_t* acc, const uint32_t* q, int q_limbs, int wide);
/* floor(2^996 / q), low 8 limbs */
void orc_barrett_mu_996(const uint32_t* q, int q_limbs, uint32_t* mu8);

#ifdef __cplusplus
}
#endif
#endif
