"""The seeded random streams of the synthetic workloads.

The reference has no generator for the benchmark distributions (its only
procedural scene is fibonacci_cluster, src/scene/synthetic.rs:6-42), so the
generator is specified here, in full, so that any implementation (numpy here,
Rust or C++ elsewhere) produces the same bytes:

* SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c):
  z += 0x9E3779B97F4A7C15; z = (z ^ z >> 30) * 0xBF58476D1CE4E5B9;
  z = (z ^ z >> 27) * 0x94D049BB133111EB; output z ^ z >> 31 (mod 2^64).
* xoshiro256** (Blackman, Vigna 2018): output rotl(s1 * 5, 7) * 9, then
  t = s1 << 17; s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3; s2 ^= t;
  s3 = rotl(s3, 45).
* One stream per (seed, domain d, index i): the SplitMix64 state starts at
  seed + 0x9E3779B97F4A7C15 * 4 * (2^32 d + i) (mod 2^64), and its next four
  outputs are s0..s3.  Each stream takes exactly four SplitMix64 steps, so
  the streams of one seed never share a SplitMix64 output.
* Uniform in [0, 1): (next >> 11) * 2^-53, in float64.
* Normal: Box-Muller in float64 on two uniforms u1, u2 taken in that order:
  r = sqrt(-2 ln(1 - u1)), z0 = r cos(2 pi u2), z1 = r sin(2 pi u2); normals
  come in pairs (z0 first), a stream's k normals use ceil(k/2) pairs, and the
  last z1 of an odd k is dropped.
* Values are computed in float64 and cast to float32 where stored.

Each workload's use of the streams (domains, draw order) is stated in the
docstring of the function that draws them.
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
_SM1, _SM2 = 0xBF58476D1CE4E5B9, 0x94D049BB133111EB

# domains of the workloads' streams
D_SCENE_GLOBAL = 0  # per cluster of a cluster-mixture scene
D_SPLAT = 1         # per splat
D_VIEW = 2          # per view of a view batch
D_GRID = 3          # per vertex of a jittered tet grid
D_IMPULSE = 4       # per vertex of a cage impulse
D_ICOSPHERE = 5     # per vertex of a noisy icosphere
D_SAMPLE = 6        # per candidate point of rejection sampling
D_PERTURB = 7       # per vertex of a cage perturbation


def splitmix64_next(z: np.ndarray):
    """One SplitMix64 step on an array of states: (new state, output)."""
    with np.errstate(over="ignore"):
        z = z + np.uint64(GOLDEN)
        x = z
        x = (x ^ (x >> np.uint64(30))) * np.uint64(_SM1)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(_SM2)
    return z, x ^ (x >> np.uint64(31))


class Streams:
    """Independent xoshiro256** streams (seed, domain, first + j), j < n,
    advanced together: each call draws the next value of every stream."""

    def __init__(self, seed: int, domain: int, n: int, first: int = 0, state=None):
        if state is not None:
            st = np.array(state, dtype=np.uint64).reshape(-1, 4)
            self.s = [st[:, w].copy() for w in range(4)]
            return
        idx = (np.uint64(domain) << np.uint64(32)) + np.arange(first, first + n, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(seed & M64) + np.uint64((GOLDEN * 4) & M64) * idx
        self.s = []
        for _ in range(4):
            z, w = splitmix64_next(z)
            self.s.append(w)
        self._t = np.empty_like(self.s[0])

    @property
    def n(self) -> int:
        return self.s[0].shape[0]

    def next_u64(self) -> np.ndarray:
        s0, s1, s2, s3 = self.s
        t = self._t if hasattr(self, "_t") else np.empty_like(s0)
        with np.errstate(over="ignore"):
            out = s1 * np.uint64(5)
            out = ((out << np.uint64(7)) | (out >> np.uint64(57))) * np.uint64(9)
        np.left_shift(s1, np.uint64(17), out=t)
        np.bitwise_xor(s2, s0, out=s2)
        np.bitwise_xor(s3, s1, out=s3)
        np.bitwise_xor(s1, s2, out=s1)
        np.bitwise_xor(s0, s3, out=s0)
        np.bitwise_xor(s2, t, out=s2)
        np.right_shift(s3, np.uint64(19), out=t)
        np.left_shift(s3, np.uint64(45), out=s3)
        np.bitwise_or(s3, t, out=s3)
        return out

    def _unit(self) -> np.ndarray:
        return (self.next_u64() >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def uniform(self, k: int = 1) -> np.ndarray:
        """[n, k] float64 in [0, 1), k draws per stream in order."""
        out = np.empty((k, self.n), dtype=np.float64)
        for j in range(k):
            out[j] = self._unit()
        return out.T

    def normal(self, k: int) -> np.ndarray:
        """[n, k] float64 standard normals (Box-Muller pairs, z0 first)."""
        pairs = (k + 1) // 2
        out = np.empty((2 * pairs, self.n), dtype=np.float64)
        for p in range(pairs):
            u1 = self._unit()
            u2 = self._unit()
            r = np.sqrt(-2.0 * np.log(1.0 - u1))
            a = 2.0 * np.pi * u2
            np.multiply(r, np.cos(a), out=out[2 * p])
            np.multiply(r, np.sin(a), out=out[2 * p + 1])
        return out[:k].T


def chunked(seed: int, domain: int, n: int, fn, first: int = 0, chunk: int = 1 << 16):
    """fn(streams) on the streams first..first+n-1 in chunks of `chunk`
    streams, on a thread pool (numpy releases the GIL); the chunks' results
    (arrays with one row per stream) concatenated in stream order."""
    from concurrent.futures import ThreadPoolExecutor
    import os

    starts = list(range(0, n, chunk))
    if len(starts) <= 1:
        return fn(Streams(seed, domain, n, first))
    workers = max(1, min(16, os.cpu_count() or 1, len(starts)))
    with ThreadPoolExecutor(workers) as ex:
        parts = list(ex.map(lambda c0: fn(Streams(seed, domain, min(chunk, n - c0), first + c0)),
                            starts))
    if isinstance(parts[0], tuple):
        return tuple(np.concatenate([p[i] for p in parts]) for i in range(len(parts[0])))
    return np.concatenate(parts)


def uniform(seed: int, domain: int, n: int, k: int, lo=0.0, hi=1.0, first: int = 0) -> np.ndarray:
    """[n, k]: k uniforms in [lo, hi) from each of the streams first..first+n-1."""
    u = chunked(seed, domain, n, lambda st: st.uniform(k), first)
    return np.asarray(lo) + (np.asarray(hi) - np.asarray(lo)) * u
