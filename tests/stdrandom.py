"""Restatements, from their published definitions, of the standard-library random pieces the
reference's Omega is built of (test infrastructure, used to pin the oracle's Omega stream).

* MT19937: Matsumoto & Nishimura, "Mersenne Twister: a 623-dimensionally equidistributed uniform
  pseudo-random number generator", ACM TOMACS 8(1), 1998 (w=32, n=624, m=397, r=31,
  a=0x9908B0DF, tempering u=11, s=7/b=0x9D2C5680, t=15/c=0xEFC60000, l=18, initialization
  multiplier f=1812433253), which is what C++ [rand.predef] names std::mt19937.
* seed_seq::generate: the algorithm of C++ [rand.util.seedseq] (constants 0x8b8b8b8b, 1664525,
  1566083941; T(x) = x ^ (x >> 27)).
* generate_canonical<double, 53> over a 32-bit engine: ceil(53 / 32) = 2 draws,
  (g0 + g1 * 2^32) / 2^64 in double arithmetic, values >= 1 replaced by the largest double
  below 1 ([rand.util.canonical]; libstdc++ header bits/random.tcc).
* normal_distribution<double>: the Marsaglia-Bray polar method (Marsaglia & Bray, SIAM Review
  6(3), 1964) as libstdc++ states it: x = 2u - 1, y = 2v - 1, reject r2 = x^2 + y^2 outside
  (0, 1], mult = sqrt(-2 log(r2) / r2), return y * mult and keep x * mult for the next call.

All arithmetic on doubles goes through Python floats and `math` (the C library's log and sqrt),
one operation at a time, so every rounding is that of the C++ expression it restates.
"""
import math

import numpy as np

M32 = 0xFFFFFFFF


class MT19937:
    N, M = 624, 397
    MATRIX_A, UPPER, LOWER = 0x9908B0DF, 0x80000000, 0x7FFFFFFF

    def __init__(self, seed):
        x = [seed & M32]
        for i in range(1, self.N):
            prev = x[-1]
            x.append((1812433253 * (prev ^ (prev >> 30)) + i) & M32)
        self.state = np.array(x, dtype=np.uint64)
        self.buf = None
        self.pos = self.N

    def _twist(self):
        s = self.state
        n, m = self.N, self.M
        out = s.copy()
        for i in range(n):  # in place, in order: entries past i-m already use updated words
            y = (int(out[i]) & self.UPPER) | (int(out[(i + 1) % n]) & self.LOWER)
            v = int(out[(i + m) % n]) ^ (y >> 1)
            if y & 1:
                v ^= self.MATRIX_A
            out[i] = v
        self.state = out
        y = out.copy()
        y ^= y >> np.uint64(11)
        y ^= (y << np.uint64(7)) & np.uint64(0x9D2C5680)
        y ^= (y << np.uint64(15)) & np.uint64(0xEFC60000)
        y ^= y >> np.uint64(18)
        self.buf = (y & np.uint64(M32)).astype(np.uint32)
        self.pos = 0

    def __call__(self):
        if self.pos >= self.N:
            self._twist()
        v = int(self.buf[self.pos])
        self.pos += 1
        return v

    def draws(self, count):
        return np.array([self() for _ in range(count)], dtype=np.uint32)


def seed_seq_generate(words, count):
    """std::seed_seq(words).generate(begin, begin + count)."""
    v = [int(w) & M32 for w in words]
    s, n = len(v), int(count)
    if n == 0:
        return np.zeros(0, np.uint32)
    b = [0x8B8B8B8B] * n
    t = 11 if n >= 623 else 7 if n >= 68 else 5 if n >= 39 else 3 if n >= 7 else (n - 1) // 2
    p = (n - t) // 2
    q = p + t
    m = max(s + 1, n)

    def T(x):
        return x ^ (x >> 27)

    for k in range(m):
        r1 = (1664525 * T(b[k % n] ^ b[(k + p) % n] ^ b[(k - 1) % n])) & M32
        if k == 0:
            r2 = r1 + s
        elif k <= s:
            r2 = r1 + k % n + v[k - 1]
        else:
            r2 = r1 + k % n
        r2 &= M32
        b[(k + p) % n] = (b[(k + p) % n] + r1) & M32
        b[(k + q) % n] = (b[(k + q) % n] + r2) & M32
        b[k % n] = r2
    for k in range(m, m + n):
        r3 = (1566083941 * T((b[k % n] + b[(k + p) % n] + b[(k - 1) % n]) & M32)) & M32
        r4 = (r3 - k % n) & M32
        b[(k + p) % n] ^= r3
        b[(k + q) % n] ^= r4
        b[k % n] = r4
    return np.array(b, dtype=np.uint32)


def canonical53(gen, draws=2):
    """generate_canonical<double, 53> over a 32-bit engine (`draws` = 2 for doubles)."""
    total, scale = 0.0, 1.0
    for _ in range(draws):
        total += float(gen()) * scale
        scale *= 4294967296.0
    u = total / scale
    if u >= 1.0:
        u = math.nextafter(1.0, 0.0)
    return u


class NormalPolar:
    """normal_distribution<double>(0, 1) as the polar method with one value kept in reserve."""

    def __init__(self, gen, draws=2):
        self.gen, self.draws, self.saved = gen, draws, None

    def __call__(self):
        if self.saved is not None:
            v, self.saved = self.saved, None
            return v * 1.0 + 0.0
        while True:
            x = 2.0 * canonical53(self.gen, self.draws) - 1.0
            y = 2.0 * canonical53(self.gen, self.draws) - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        self.saved = x * mult
        return y * mult * 1.0 + 0.0


def normals(seed, count, draws=2):
    dist = NormalPolar(MT19937(seed), draws)
    return np.array([dist() for _ in range(count)])


def omega_chain(seed, length, col_block=True, nprocs=1, rank=0, threads=1):
    """The reference's Omega vector for one rank: the seed chain (col-block: seed_seq{seed} over
    nprocs words, this rank's word; row-block: seed itself), seed_seq{s0} over `threads` words,
    thread t filling chunk t (len/threads values, the last thread the remainder)."""
    s0 = seed & M32
    if col_block:
        s0 = int(seed_seq_generate([s0], nprocs)[rank])
    per_thread = seed_seq_generate([s0], threads)
    out = np.empty(length)
    chunk = length // threads
    for t in range(threads):
        start = chunk * t
        cnt = length - start if t == threads - 1 else chunk
        out[start:start + cnt] = normals(int(per_thread[t]), cnt)
    return out
