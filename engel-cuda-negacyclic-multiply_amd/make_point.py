#!/usr/bin/env python3
"""Write a point directory for the polymul/negacyclic harness protocol.

Layout (the one the platform's generated harness reads, main.cpp:47-87):
  <dir>/manifest.json          {"N": N, "W": bits(q), "L": L, "cases": C}
  <dir>/point/q.bin            L little-endian u32 limbs of q
  <dir>/cases/NNNNNN/a.bin     N*L u32 (coefficient i, limb l at i*L + l)
  <dir>/cases/NNNNNN/b.bin
Random cases use the library's seeded generator (nc_gen_uniform: SplitMix64,
a from seed 2s, b from seed 2s+1, uniform in [0, q)).  --edges prepends the
three edge cases of the benchmark configuration: every coefficient q-1,
all zero, and v X^(N-1) times X (the negacyclic sign wrap).

  python make_point.py OUT_DIR --N 16384 --cases 4 [--q hex] [--seed 1] [--edges]
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent))
import ncmul  # noqa: E402

Q868 = (1 << 868) - 467


def limbs(x: int, L: int) -> np.ndarray:
    return np.frombuffer(int(x).to_bytes(4 * L, "little"), dtype="<u4").astype(np.uint32)


def edge_cases(N: int, q: int, L: int):
    full = np.tile(limbs(q - 1, L), (N, 1))
    zero = np.zeros((N, L), dtype=np.uint32)
    a = zero.copy()
    b = zero.copy()
    a[N - 1] = limbs((q * 3) // 7, L)
    b[1 % N] = limbs(1, L)
    return [(full, full.copy()), (zero, full.copy()), (a, b)]


def write_point(out: Path, N: int, q: int, cases: int, seed: int, edges: bool) -> int:
    L = (q.bit_length() + 31) // 32
    qv = limbs(q, L)
    pairs = edge_cases(N, q, L) if edges else []
    for i in range(cases):
        s = seed + i
        pairs.append((ncmul.gen_uniform(2 * s, N, qv), ncmul.gen_uniform(2 * s + 1, N, qv)))
    (out / "point").mkdir(parents=True, exist_ok=True)
    (out / "point" / "q.bin").write_bytes(qv.astype("<u4").tobytes())
    for i, (a, b) in enumerate(pairs):
        d = out / "cases" / f"{i:06d}"
        d.mkdir(parents=True, exist_ok=True)
        (d / "a.bin").write_bytes(np.ascontiguousarray(a, dtype="<u4").tobytes())
        (d / "b.bin").write_bytes(np.ascontiguousarray(b, dtype="<u4").tobytes())
    (out / "manifest.json").write_text(json.dumps({"N": N, "W": q.bit_length(), "L": L, "cases": len(pairs)}))
    return len(pairs)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("out", type=Path)
    ap.add_argument("--N", type=int, default=16384)
    ap.add_argument("--q", type=lambda s: int(s, 16), default=Q868, help="modulus in hex (default 2^868-467)")
    ap.add_argument("--cases", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--edges", action="store_true")
    args = ap.parse_args()
    n = write_point(args.out, args.N, args.q, args.cases, args.seed, args.edges)
    print(f"wrote {n} cases to {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
