"""CPU tests that pin the oracle (oracle/nc_oracle.c) before anything trusts it.

Pins, chosen so a plausible slip (dropped term, wrong sign or index,
transposed operand, off-by-one prime) fails at least one of them:
  * golden vectors from pure Python ints (tests/golden, Kronecker
    substitution) for every N from 2 to 65536 and three moduli;
  * closed forms of the reference's edge inputs (all q-1, zeros, the
    X^(N-1) * X sign wrap, monomials, b = 1, b = q-1);
  * algebraic invariants (commutativity, bilinearity, X^N = -1);
  * word primitives against exact Python-int arithmetic;
  * the reference's own init-time known-answer tests (solve.cu:1151-1166,
    1285-1314, 1370-1415) via orc_selftest;
  * stage values (residues, product residues, CRT quotient decision) against
    exact Python-int recomputation.
"""
from __future__ import annotations

import hashlib
import random

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN_DIR, modulus

SMALL_SCHOOLBOOK_MAX = 2048


def sha(x: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(x, dtype="<u4").tobytes()).hexdigest()


def golden_cases():
    from conftest import load_golden

    return load_golden()["cases"]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: f"{c['modulus']}-N{c['N']}-s{c['seed']}")
def test_golden_restated_path(case):
    """Generator (C) reproduces the Python-int inputs; the restated reference
    RNS path reproduces the Python-int product bit for bit."""
    q = modulus(case["modulus"])
    N, s = case["N"], case["seed"]
    a = O.gen_poly(2 * s, N, q)
    b = O.gen_poly(2 * s + 1, N, q)
    assert sha(a) == case["sha256_a"]
    assert sha(b) == case["sha256_b"]
    plan = O.Plan(N, q)
    # the reference's init accepts q868 and q200 and rejects q868b at its
    # q-Barrett check (solve.cu:1396-1415, two corrections only)
    assert plan.selftest() == (8 if case["modulus"] == "q868b" else 0)
    c = plan.multiply(a, b)
    assert sha(c) == case["sha256_c"]
    for idx, hx in case["samples"].items():
        assert O.limbs_to_int(c[int(idx)]) == int(hx, 16)
    if "c_file" in case:
        ref = np.fromfile(GOLDEN_DIR / case["c_file"], dtype="<u4").reshape(N, -1)
        assert np.array_equal(c, ref)


@pytest.mark.parametrize("case", [c for c in golden_cases() if c["N"] <= SMALL_SCHOOLBOOK_MAX],
                         ids=lambda c: f"{c['modulus']}-N{c['N']}")
def test_golden_schoolbook(case):
    q = modulus(case["modulus"])
    N, s = case["N"], case["seed"]
    a, b = O.gen_poly(2 * s, N, q), O.gen_poly(2 * s + 1, N, q)
    assert sha(O.schoolbook(a, b, q)) == case["sha256_c"]


@pytest.mark.parametrize("case", [c for c in golden_cases() if c["N"] > SMALL_SCHOOLBOOK_MAX],
                         ids=lambda c: f"{c['modulus']}-N{c['N']}")
def test_golden_schoolbook_samples(case):
    q = modulus(case["modulus"])
    N, s = case["N"], case["seed"]
    a, b = O.gen_poly(2 * s, N, q), O.gen_poly(2 * s + 1, N, q)
    idx = sorted(int(i) for i in case["samples"])[:6]
    got = O.schoolbook_coeffs(a, b, q, idx)
    for row, i in zip(got, idx):
        assert O.limbs_to_int(row) == int(case["samples"][str(i)], 16)


@pytest.mark.parametrize("N", [8, 1024, 16384])
def test_garner_matches_fast_crt(N, q868):
    a, b = O.gen_poly(100, N, q868), O.gen_poly(101, N, q868)
    plan = O.Plan(N, q868)
    res = plan.product_residues(a, b)
    assert np.array_equal(plan.crt_garner(res), plan.crt_fast(res))


# ---------------------------------------------------------------------------
# closed forms (edge inputs: all q - 1, zero, the X^N = -1 sign wrap, monomials)
# ---------------------------------------------------------------------------

def both(a, b, q):
    N = a.shape[0]
    plan = O.Plan(N, q)
    c1 = plan.multiply(a, b)
    if N <= 1024:
        assert np.array_equal(c1, O.schoolbook(a, b, q))
    return c1


@pytest.mark.parametrize("N", [2, 16, 1024, 16384])
def test_all_q_minus_one(N, q868):
    qi = O.limbs_to_int(q868)
    a = np.tile(O.int_to_limbs(qi - 1, 28), (N, 1))
    c = both(a, a.copy(), q868)
    for k in [0, 1, N // 2 - 1, N // 2, N - 1]:
        assert O.limbs_to_int(c[k]) == (2 * k + 2 - N) % qi


@pytest.mark.parametrize("N", [4, 16384])
def test_zero(N, q868):
    z = np.zeros((N, 28), dtype=np.uint32)
    a = O.gen_poly(5, N, q868)
    assert not both(z, a, q868).any()
    assert not both(a, z, q868).any()


@pytest.mark.parametrize("N", [2, 1024, 16384])
def test_sign_wrap(N, q868):
    """a = v X^(N-1), b = X  =>  c = -v (coefficient 0), zero elsewhere."""
    qi = O.limbs_to_int(q868)
    v = (qi * 3) // 7
    a = np.zeros((N, 28), dtype=np.uint32)
    b = np.zeros((N, 28), dtype=np.uint32)
    a[N - 1] = O.int_to_limbs(v, 28)
    b[1] = O.int_to_limbs(1, 28)
    c = both(a, b, q868)
    assert O.limbs_to_int(c[0]) == qi - v
    assert not c[1:].any()


def test_monomials(q868):
    qi = O.limbs_to_int(q868)
    N = 64
    rng = random.Random(7)
    for _ in range(6):
        i, j = rng.randrange(N), rng.randrange(N)
        u, w = rng.randrange(qi), rng.randrange(qi)
        a = np.zeros((N, 28), dtype=np.uint32)
        b = np.zeros((N, 28), dtype=np.uint32)
        a[i], b[j] = O.int_to_limbs(u, 28), O.int_to_limbs(w, 28)
        c = both(a, b, q868)
        k = (i + j) % N
        want = (u * w) % qi if i + j < N else (-u * w) % qi
        assert O.limbs_to_int(c[k]) == want
        c[k] = 0
        assert not c.any()


@pytest.mark.parametrize("N", [32, 16384])
def test_identity_and_negation(N, q868):
    qi = O.limbs_to_int(q868)
    a = O.gen_poly(9, N, q868)
    one = np.zeros_like(a)
    one[0] = O.int_to_limbs(1, 28)
    assert np.array_equal(both(a, one, q868), a)
    neg = np.zeros_like(a)
    neg[0] = O.int_to_limbs(qi - 1, 28)
    c = both(a, neg, q868)
    ai, ci = O.poly_to_ints(a), O.poly_to_ints(c)
    assert all(x == (qi - y) % qi for x, y in zip(ci, ai))


def test_commutative_bilinear(q868):
    qi = O.limbs_to_int(q868)
    N = 4096
    plan = O.Plan(N, q868)
    a, a2, b = O.gen_poly(11, N, q868), O.gen_poly(12, N, q868), O.gen_poly(13, N, q868)
    ab, ba = plan.multiply(a, b), plan.multiply(b, a)
    assert np.array_equal(ab, ba)
    s = O.ints_to_poly([(x + y) % qi for x, y in zip(O.poly_to_ints(a), O.poly_to_ints(a2))], 28)
    lhs = O.poly_to_ints(plan.multiply(s, b))
    r1, r2 = O.poly_to_ints(ab), O.poly_to_ints(plan.multiply(a2, b))
    assert lhs == [(x + y) % qi for x, y in zip(r1, r2)]


# ---------------------------------------------------------------------------
# word primitives vs exact Python ints
# ---------------------------------------------------------------------------

def test_selftests_kat():
    assert O.lib().orc_selftest(None) == 0
    q = modulus("q868")
    for N in (2, 1024, 16384, 65536):
        assert O.Plan(N, q).selftest() == 0


def word_cases(p):
    rng = random.Random(p)
    vals = [0, 1, 2, p - 1, p - 2, p // 2, (1 << 32) - 1, 1 << 32]
    vals += [rng.randrange(p) for _ in range(40)]
    return vals


def test_word_primitives():
    L = O.lib()
    R = 1 << 64
    q = modulus("q868")
    for p in [int(x) for x in O.Plan(16384, q).primes()[:4]] + [11, 0x3fffffffffffffa9]:
        np_ = (-pow(p, -1, R)) % R
        assert L.orc_inv_mod_2_64(p) * p % R == 1
        mag = (1 << 128) // p
        mh, ml = mag >> 64, mag & (R - 1)
        mu = L.orc_barrett_mu(p)
        assert mu == ((1 << 96) // p) % R
        for a in word_cases(p):
            for b in word_cases(p ^ 0x55)[:10]:
                b %= p
                assert L.orc_mul_mod(a, b, p) == a * b % p
                assert L.orc_mont_mul(a, b, p, np_) == a * b * pow(R, -1, p) % p
                ws = (b << 64) // p
                assert L.orc_shoup_mul(a, b, ws, p) == a * b % p
            x = a * (1 << 36) + 12345
            if x < (1 << 100):
                assert L.orc_reduce_128(x >> 64, x & (R - 1), p, mh, ml) == x % p
            if p > (1 << 32):
                # barrett_reduce's domain in the reference (Horner step of
                # solve.cu:519-525): 62-bit p, v = r 2^32 + word < p 2^32
                y = (a % p) * (1 << 32) + (a & 0xFFFFFFFF)
                assert L.orc_barrett_reduce(y >> 64, y & (R - 1), p, mu) == y % p
        assert L.orc_pow_mod(3, p - 1, p) == 1 if p != 3 else True


def test_mod_q_routines():
    q = modulus("q868")
    qi = O.limbs_to_int(q)
    mu = np.zeros(8, dtype=np.uint32)
    O.lib().orc_barrett_mu_996(O._p32(q), 28, O._p32(mu))
    assert O.limbs_to_int(mu) == (1 << 996) // qi
    rng = random.Random(3)
    for _ in range(200):
        x = rng.randrange(1 << rng.choice([800, 868, 900, 930, 934]))
        v = O.int_to_limbs(x, 34)
        O.lib().orc_mod_q_barrett(O._p32(v), O._p32(q), O._p32(mu))
        assert O.limbs_to_int(v) == x % qi
        w = O.int_to_limbs(x, 40)
        O.lib().orc_mod_q(O._p32(w), O._p32(q), 28, 40)
        assert O.limbs_to_int(w) == x % qi


def test_prime_rule_facts():
    """Prime rule of solve.cu:1179-1204 with the 868-bit q: k = 29 primes for
    N = 2^10..2^16, each p = 1 mod 2N, p < 2^62, prime, and prod p beyond the
    coefficient bound; first and last prime at N = 2^14 as recomputed
    independently (0x3fffffffffff0001, 0x3ffffffffed30001)."""
    from gen_golden import is_probable_prime

    q = modulus("q868")
    for logn in range(10, 17):
        N = 1 << logn
        pr = [int(x) for x in O.Plan(N, q).primes()]
        assert len(pr) == 29
        assert all(p % (2 * N) == 1 and p < (1 << 62) and is_probable_prime(p) for p in pr)
        assert pr == sorted(pr, reverse=True)
        M = 1
        for p in pr:
            M *= p
        assert M.bit_length() >= logn + 2 * 868 + 4
        assert (M // pr[-1]).bit_length() < logn + 2 * 868 + 4
    pr = O.Plan(16384, q).primes()
    assert int(pr[0]) == 0x3FFFFFFFFFFF0001 and int(pr[-1]) == 0x3FFFFFFFFED30001


# ---------------------------------------------------------------------------
# stage values vs exact Python ints
# ---------------------------------------------------------------------------

def test_stage_values_exact():
    q = modulus("q868")
    qi = O.limbs_to_int(q)
    N = 64
    a, b = O.gen_poly(40, N, q), O.gen_poly(41, N, q)
    a[3] = O.int_to_limbs(qi - 1, 28)  # force some negative coefficients / large values
    plan = O.Plan(N, q)
    pr = [int(x) for x in plan.primes()]
    ai, bi = O.poly_to_ints(a), O.poly_to_ints(b)
    res = plan.residues(a)
    for j, p in enumerate(pr):
        assert [int(x) for x in res[j]] == [x % p for x in ai]
    cint = []
    for k in range(N):
        s = 0
        for i in range(N):
            s += ai[i] * bi[k - i] if k - i >= 0 else -ai[i] * bi[k - i + N]
        cint.append(s)
    pres = plan.product_residues(a, b)
    for j, p in enumerate(pr):
        assert [int(x) for x in pres[j]] == [x % p for x in cint]
    M = 1
    for p in pr:
        M *= p
    hi, neg = plan.crt_decisions(pres)
    for i, c in enumerate(cint):
        X = c % M
        v = [int(pres[j][i]) * pow(M // p, -1, p) % p for j, p in enumerate(pr)]
        kappa = (sum(vj * (M // p) for vj, p in zip(v, pr)) - X) // M
        assert int(hi[i]) + int(neg[i]) == kappa + (1 if c < 0 else 0)
    assert O.poly_to_ints(plan.crt_fast(pres)) == [c % qi for c in cint]
    assert O.poly_to_ints(plan.crt_garner(pres)) == [c % qi for c in cint]
