"""The inner NTT's merged levels (merge_product30 in nc_kernels.hip): the last
MG forward stages of a and b, their pointwise product and the first MG
inverse stages run as one product mod X^(2^MG) - C per group of 2^MG
positions.  CPU checks of the two facts the kernel relies on, chosen so that
a wrong sign, twiddle index or scale fails them:

- the algebra, exactly: a merged-psi Cooley-Tukey forward transform stopped
  above stage MG, the group products with C = +/- (stage-MG twiddle), the
  Gentleman-Sande inverse from stage MG and the final scale N^-1 2^MG give
  the schoolbook negacyclic product (N = 8 .. 64);
- the word arithmetic, as the kernel performs it (u32 wrap-around, 64-bit
  sums, the forward twiddle stored as 2^32 - w with its Shoup quotient):
  for inputs anywhere in [0, 4p), up to 4p - 1, and the largest primes
  p = 1 mod 2^17 below 2^30, no 64-bit sum overflows, every output is below
  2p, and it equals the exact group product times 2^-32 mod p.
"""
from __future__ import annotations

import random

import pytest

M32 = (1 << 32) - 1


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if a % n == 0:
            continue
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def largest_primes(count: int, step: int = 1 << 17, below: int = 1 << 30) -> list[int]:
    out, p = [], (below - 1) // step * step + 1
    while len(out) < count:
        if p < below and is_prime(p):
            out.append(p)
        p -= step
    return out


def bitrev(x: int, bits: int) -> int:
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def negacyclic(a, b, p):
    n = len(a)
    c = [0] * n
    for i in range(n):
        for j in range(n):
            if i + j < n:
                c[i + j] = (c[i + j] + a[i] * b[j]) % p
            else:
                c[i + j - n] = (c[i + j - n] - a[i] * b[j]) % p
    return c


def psi_table(n: int, p: int) -> list[int]:
    """tw[x] = psi^bitrev(x) for a primitive 2n-th root psi (stage s, group i: tw[2^s + i])."""
    logn = n.bit_length() - 1
    for g in range(2, p):
        psi = pow(g, (p - 1) // (2 * n), p)
        if pow(psi, n, p) == p - 1:
            return [pow(psi, bitrev(x, logn), p) for x in range(n)]
    raise AssertionError("no root")


def tw_of(tw, logn, beta, x):
    return tw[(1 << (logn - 1 - beta)) + (x >> (beta + 1))]


def forward(a, tw, p, stop):
    a, n = a[:], len(a)
    logn = n.bit_length() - 1
    for beta in range(logn - 1, stop - 1, -1):
        span = 1 << beta
        for x in range(n):
            if not x & span:
                u, v = a[x], a[x + span] * tw_of(tw, logn, beta, x) % p
                a[x], a[x + span] = (u + v) % p, (u - v) % p
    return a


def inverse(a, tw, p, start):
    a, n = a[:], len(a)
    logn = n.bit_length() - 1
    for beta in range(start, logn):
        span = 1 << beta
        for x in range(n):
            if not x & span:
                winv = pow(tw_of(tw, logn, beta, x), p - 2, p)
                u, v = a[x], a[x + span]
                a[x], a[x + span] = (u + v) % p, (u - v) * winv % p
    return a


def group_product(A, B, x0, g, C, p):
    out = []
    for k in range(g):
        acc = 0
        for i in range(g):
            j = k - i
            acc += A[x0 + i] * (B[x0 + j] if j >= 0 else B[x0 + j + g] * C)
        out.append(acc % p)
    return out


@pytest.mark.parametrize("n", [8, 16, 32, 64])
@pytest.mark.parametrize("mg", [1, 2])
def test_merged_levels_give_the_negacyclic_product(n, mg):
    p = largest_primes(1)[0]
    tw = psi_table(n, p)
    logn = n.bit_length() - 1
    rng = random.Random(n * 10 + mg)
    a = [rng.randrange(p) for _ in range(n)]
    b = [rng.randrange(p) for _ in range(n)]
    A, B = forward(a, tw, p, mg), forward(b, tw, p, mg)
    g = 1 << mg
    prod = [0] * n
    for x0 in range(0, n, g):
        t = tw_of(tw, logn, mg, x0)
        C = (p - t) if (x0 >> mg) & 1 else t
        # C is the square of the group's stage-(mg - 1) twiddle
        assert tw_of(tw, logn, mg - 1, x0) ** 2 % p == C
        prod[x0:x0 + g] = group_product(A, B, x0, g, C, p)
    got = [v * pow(n, p - 2, p) * (1 << mg) % p for v in inverse(prod, tw, p, mg)]
    assert got == negacyclic(a, b, p)
    # the same product without the merge: the scale differs by 2^mg exactly
    A0, B0 = forward(a, tw, p, 0), forward(b, tw, p, 0)
    sep = inverse([x * y % p for x, y in zip(A0, B0)], tw, p, 0)
    assert [v * pow(n, p - 2, p) % p for v in sep] == got


def merge_words(va, vb, mt, p, mg):
    """merge_product30 on 16 registers, u32 / u64 as the kernel computes."""
    g, p2, mneg = 1 << mg, 2 * p, (-pow(p, -1, 1 << 32)) & M32

    def sub_if_ge(x, m):
        return min(x, (x - m) & M32)

    out = [0] * 16
    for g0 in range(0, 16, g):
        w, ws = mt[g0 >> (mg + 1)]
        a = [sub_if_ge(sub_if_ge(va[g0 + i], p2), p) for i in range(g)]
        b = [sub_if_ge(vb[g0 + i], p2) for i in range(g)]
        bc = [0] * g
        for j in range(1, g):
            y = vb[g0 + j]
            nv = (((y * ws) >> 32) * p + y * w) & M32
            bc[j] = (nv + p2) & M32 if (g0 >> mg) & 1 else (-nv) & M32
        for k in range(g):
            t = sum(a[i] * (b[k - i] if k - i >= 0 else bc[k - i + g]) for i in range(g))
            m = (t * mneg) & M32
            assert t + m * p < 1 << 64
            r = (t + m * p) >> 32
            out[g0 + k] = r if mg == 1 else sub_if_ge(r, p2)
    return out


@pytest.mark.parametrize("mg", [1, 2])
def test_merged_word_arithmetic_bounds_and_value(mg):
    rng = random.Random(100 + mg)
    for p in largest_primes(3):
        rinv = pow(1 << 32, -1, p)
        for trial in range(400):
            if trial == 0:
                va = vb = [4 * p - 1] * 16
            elif trial % 3 == 1:
                va = [rng.randrange(3 * p, 4 * p) for _ in range(16)]
                vb = [rng.randrange(3 * p, 4 * p) for _ in range(16)]
            else:
                va = [rng.randrange(4 * p) for _ in range(16)]
                vb = [rng.randrange(4 * p) for _ in range(16)]
            ts = [rng.randrange(1, p) for _ in range(8 >> mg)]
            mt = [((1 << 32) - t, (t << 32) // p) for t in ts]
            got = merge_words(va, vb, mt, p, mg)
            assert max(got) < 2 * p
            g = 1 << mg
            for g0 in range(0, 16, g):
                t = ts[g0 >> (mg + 1)]
                C = (p - t) if (g0 >> mg) & 1 else t
                want = group_product(va, vb, g0, g, C, p)
                assert [x % p for x in got[g0:g0 + g]] == [w * rinv % p for w in want], (p, mg, trial, g0)
