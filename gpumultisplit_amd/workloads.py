"""Seeded synthetic inputs for the multisplit workloads (numpy, host side).

Distributions follow the reference's generators
(src/random_generator.cpp:43-215) in meaning, with our own seeded PRNG
(numpy PCG64) instead of the reference's time-seeded rand():

* ``uniform_keys``: keys uniform over [0, 2^32).
* ``bucket_keys``: a bucket id per key, then a key inside that bucket's range
  of the range identifier, as ``single_rnd_generator`` mode 0 puts the bucket
  id in the top log2(B) bits (random_generator.cpp:43-51).
* ``delta_keys``: key uniform in [b*delta, (b+1)*delta) for a uniform bucket b,
  as ``single_rnd_generator`` mode 1 (random_generator.cpp:52-57).
* ``zipf_bucket_ids``: bucket ids with P(b) proportional to 1/(b+1)^s.
* ``onehot_bucket_ids``: a share ``1-alpha`` of keys in one random bucket and
  the rest uniform, shuffled (hockey-stick, random_generator.cpp:140-194).
* ``permutation_keys``: a random permutation of 0..n-1
  (random_permute, random_generator.cpp:23-40; main_compaction.cu input).
"""
from __future__ import annotations

import math

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def uniform_keys(n: int, seed: int) -> np.ndarray:
    return rng(seed).integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)


def zipf_bucket_ids(n: int, num_buckets: int, seed: int, s: float = 1.1) -> np.ndarray:
    w = 1.0 / np.power(np.arange(1, num_buckets + 1, dtype=np.float64), s)
    w /= w.sum()
    return rng(seed).choice(num_buckets, size=n, p=w).astype(np.uint32)


def onehot_bucket_ids(n: int, num_buckets: int, seed: int, hot_share: float = 0.9) -> np.ndarray:
    g = rng(seed)
    n_hot = int(round(hot_share * n))
    hot = int(g.integers(0, num_buckets))
    ids = np.concatenate([
        g.integers(0, num_buckets, size=n - n_hot, dtype=np.int64),
        np.full(n_hot, hot, dtype=np.int64),
    ]).astype(np.uint32)
    g.shuffle(ids)
    return ids


def bucket_keys(ids: np.ndarray, num_buckets: int, seed: int) -> np.ndarray:
    """Keys whose range-identifier bucket, floor(key*B/2^32), equals ``ids``."""
    ids = ids.astype(np.uint64)
    B = np.uint64(num_buckets)
    lo = (ids << np.uint64(32)) + B - np.uint64(1)
    lo //= B                                    # ceil(b * 2^32 / B)
    hi = ((ids + np.uint64(1)) << np.uint64(32)) + B - np.uint64(1)
    hi //= B                                    # ceil((b+1) * 2^32 / B)
    u = rng(seed).random(size=ids.size)
    span = (hi - lo).astype(np.float64)
    k = lo + np.minimum(np.floor(u * span).astype(np.uint64), (hi - lo) - np.uint64(1))
    return k.astype(np.uint32)


def delta_keys(n: int, num_buckets: int, delta: int, seed: int) -> np.ndarray:
    g = rng(seed)
    b = g.integers(0, num_buckets, size=n, dtype=np.uint64)
    off = g.integers(0, delta, size=n, dtype=np.uint64)
    k = b * np.uint64(delta) + off
    if k.size and int(k.max()) >= (1 << 32):
        raise ValueError("delta * num_buckets exceeds the 32-bit key range")
    return k.astype(np.uint32)


def permutation_keys(n: int, seed: int) -> np.ndarray:
    return rng(seed).permutation(n).astype(np.uint32)


def keys_for(dist: str, n: int, num_buckets: int, seed: int) -> np.ndarray:
    """Keys for the plain-multisplit sweep under the range identifier."""
    if dist == "uniform":
        return uniform_keys(n, seed)
    if dist == "zipf":
        return bucket_keys(zipf_bucket_ids(n, num_buckets, seed), num_buckets, seed + 1)
    if dist == "onehot90":
        return bucket_keys(onehot_bucket_ids(n, num_buckets, seed), num_buckets, seed + 1)
    raise ValueError(f"unknown distribution {dist!r}")


def log_uniform_sizes(count: int, lo_log2: float, hi_log2: float, seed: int) -> np.ndarray:
    """Sizes round(2^U), U ~ Uniform[lo_log2, hi_log2]."""
    u = rng(seed).uniform(lo_log2, hi_log2, size=count)
    return np.rint(np.power(2.0, u)).astype(np.int64)


def ceil_log2(x: int) -> int:
    return 0 if x <= 1 else int(math.ceil(math.log2(x)))
