#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ from pure Python integers.

Nothing here runs or imports the reference (it has no Python and its CUDA
cannot build here); the vectors come from the mathematical definition

    c_k = sum_{i+j=k} a_i b_j - sum_{i+j=k+N} a_i b_j   (mod q)

evaluated by Kronecker substitution on Python ints (one big-integer product
per case), which shares no code with the C oracle or the GPU library.

Moduli (the reference reads q at run time from point/q.bin, main.cpp:55, and
ships none, so the project fixes its own):
  q868  = 2^868 - 467, the largest 868-bit prime (L = 28 limbs, W = 868)
  q868b = a seeded random 868-bit prime (generality: no special form)
  q200  = 2^200 - 75, the largest 200-bit prime (L = 7, generic-limb path)

Inputs: SplitMix64(seed) streams; every candidate coefficient draws
ceil(L/2) u64 words (word w gives limbs 2w, 2w+1 little endian), the top limb
is masked to bits(q) - 32(L-1) bits, and candidates >= q are redrawn.
Polynomial a uses seed 2s, b uses seed 2s+1.

Output: golden.json (SHA-256 of a.bin, b.bin, c.bin in the harness's raw
little-endian u32 layout, plus sampled coefficients in hex) and one full
c.bin for N=1024.  Run: python tests/golden/gen_golden.py  (about 3 minutes;
the N=65536 case dominates).
"""
from __future__ import annotations

import hashlib
import json
import random
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
MASK64 = (1 << 64) - 1


def is_probable_prime(n: int) -> bool:
    if n < 2:
        return False
    small = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47]
    for p in small:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    rng = random.Random(n & MASK64)
    bases = small + [rng.randrange(2, n - 1) for _ in range(24)]
    for a in bases:
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


def prev_prime(n: int) -> int:
    n -= 1 if n % 2 == 0 else 2
    while not is_probable_prime(n):
        n -= 2
    return n


def next_prime(n: int) -> int:
    n += 1 if n % 2 == 0 else 2
    while not is_probable_prime(n):
        n += 2
    return n


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)


def gen_poly(seed: int, N: int, q: int, L: int) -> list:
    qbits = q.bit_length()
    top = qbits - 32 * (L - 1)
    mask_bits = 32 * (L - 1) + min(top, 32)
    words = (L + 1) // 2
    g = SplitMix64(seed)
    out = []
    while len(out) < N:
        x = 0
        for w in range(words):
            x |= g.next() << (64 * w)
        x &= (1 << mask_bits) - 1  # drops limb L (odd L) and masks the top limb
        if x < q:
            out.append(x)
    return out


def to_bytes(xs, L: int) -> bytes:
    return b"".join(int(x).to_bytes(4 * L, "little") for x in xs)


def negacyclic(a, b, q: int) -> list:
    N = len(a)
    w = ((N * q * q).bit_length() + 8) // 8 * 8
    wb = w // 8
    A = int.from_bytes(b"".join(int(x).to_bytes(wb, "little") for x in a), "little")
    B = int.from_bytes(b"".join(int(x).to_bytes(wb, "little") for x in b), "little")
    raw = (A * B).to_bytes(2 * N * wb, "little")
    lin = [int.from_bytes(raw[wb * k: wb * (k + 1)], "little") for k in range(2 * N)]
    return [(lin[k] - lin[k + N]) % q for k in range(N)]


def sample_indices(N: int, seed: int, count: int = 16) -> list:
    rng = random.Random(seed * 1000003 + N)
    base = sorted({0, 1, N // 2, N - 1} & set(range(N)))
    extra = set()
    while len(extra) < min(count, N) - len(base) and len(extra) + len(base) < N:
        i = rng.randrange(N)
        if i not in base:
            extra.add(i)
    return sorted(set(base) | extra)


def moduli() -> dict:
    q868 = prev_prime(1 << 868)
    assert q868 == (1 << 868) - 467
    g = SplitMix64(0x868)
    x = 0
    for w in range(14):
        x |= g.next() << (64 * w)
    x &= (1 << 868) - 1
    x |= (1 << 867) | 1
    q868b = next_prime(x)
    assert q868b.bit_length() == 868
    q200 = prev_prime(1 << 200)
    return {
        "q868": {"q": q868, "L": 28, "W": 868},
        "q868b": {"q": q868b, "L": 28, "W": 868},
        "q200": {"q": q200, "L": 7, "W": 200},
    }


CASES = (
    # (modulus, N, seed, write full c.bin)
    [("q868", 1 << e, 1, False) for e in range(1, 10)]
    + [("q868", 1024, 1, True), ("q868", 2048, 1, False), ("q868", 4096, 1, False), ("q868", 8192, 1, False)]
    + [("q868", 16384, 1, False), ("q868", 32768, 1, False), ("q868", 65536, 1, False)]
    + [("q868b", 1024, 3, False), ("q868b", 16384, 3, False)]
    + [("q200", 16, 5, False), ("q200", 1024, 5, False)]
)


def main() -> int:
    mods = moduli()
    doc = {
        "about": "Golden vectors for c = a*b mod (X^N+1, q); generated by gen_golden.py from Python ints.",
        "generator": "SplitMix64; a seed 2s, b seed 2s+1; ceil(L/2) u64 per candidate; top limb masked; reject >= q",
        "moduli": {k: {"hex": hex(v["q"]), "L": v["L"], "W": v["W"]} for k, v in mods.items()},
        "cases": [],
    }
    for name, m in mods.items():
        (HERE / f"{name}.bin").write_bytes(int(m["q"]).to_bytes(4 * m["L"], "little"))
    for name, N, seed, full in CASES:
        q, L = mods[name]["q"], mods[name]["L"]
        t0 = time.time()
        a = gen_poly(2 * seed, N, q, L)
        b = gen_poly(2 * seed + 1, N, q, L)
        c = negacyclic(a, b, q)
        cb = to_bytes(c, L)
        idx = sample_indices(N, seed)
        case = {
            "modulus": name, "N": N, "seed": seed,
            "sha256_a": hashlib.sha256(to_bytes(a, L)).hexdigest(),
            "sha256_b": hashlib.sha256(to_bytes(b, L)).hexdigest(),
            "sha256_c": hashlib.sha256(cb).hexdigest(),
            "samples": {str(i): hex(c[i]) for i in idx},
        }
        if full:
            fname = f"c_{name}_N{N}_s{seed}.bin"
            (HERE / fname).write_bytes(cb)
            case["c_file"] = fname
        doc["cases"].append(case)
        print(f"{name} N={N} seed={seed}: {time.time() - t0:.1f}s", flush=True)
    (HERE / "golden.json").write_text(json.dumps(doc, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
