"""Writes the golden fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

The reference ships no golden vectors, so the fixtures are produced by the C
oracle and, before anything is written, checked against an independent
computation of the same result with numpy (stable argsort by bucket id, sort,
bincount, float32 binning, partition); a mismatch aborts.  Inputs come from
seeded PCG64 streams; small cases store inputs and expected outputs (for a
multisplit or sort with values = original index, the output values are the
stable permutation and the output keys are keys[permutation]); the 2^20-key
case stores its seed and the SHA-256 of the expected output.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from gpumultisplit_amd import workloads as W  # noqa: E402


def ids_np(keys, B, kind, param):
    k = keys.astype(np.uint64)
    if kind == oracle.RANGE:
        b = (k * np.uint64(B)) >> np.uint64(32)
    elif kind == oracle.DELTA:
        b = k // np.uint64(param)
    elif kind == oracle.DIGIT:
        b = (k >> np.uint64(param)) & np.uint64(B - 1)
    elif kind == oracle.IDENTITY:
        b = k
    else:
        b = k & np.uint64((1 << param) - 1)
    return np.minimum(b, np.uint64(B - 1)).astype(np.uint32)


def multisplit_np(keys, B, kind, param, vals):
    ids = ids_np(keys, B, kind, param)
    order = np.argsort(ids, kind="stable")
    counts = np.bincount(ids, minlength=B)
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32)
    return keys[order], (vals[order] if vals is not None else None), offs


def check(a, b, what):
    if not np.array_equal(a, b):
        raise SystemExit(f"oracle and numpy disagree on {what}; nothing written")


def main():
    out = {}
    # stable multisplit cases: (name, n, B, kind, param, kv, key source)
    cases = [
        ("ms_range8_ko", 3001, 8, oracle.RANGE, 0, False, "uniform"),
        ("ms_range256_kv", 6145, 256, oracle.RANGE, 0, True, "uniform"),
        ("ms_range100_kv", 2500, 100, oracle.RANGE, 0, True, "uniform"),
        ("ms_zipf64_kv", 2000, 64, oracle.RANGE, 0, True, "zipf"),
        ("ms_delta32_kv", 3000, 32, oracle.DELTA, 1000, True, "delta"),
        ("ms_identity16_ko", 1000, 16, oracle.IDENTITY, 0, False, "small"),
        ("ms_lowbits2_kv", 6200, 2, oracle.LOWBITS, 1, True, "uniform"),
        ("ms_digit128_kv", 1500, 128, oracle.DIGIT, 25, True, "uniform"),
    ]
    for i, (name, n, B, kind, param, kv, src) in enumerate(cases):
        seed = 9000 + i
        if src == "uniform":
            keys = W.uniform_keys(n, seed)
        elif src == "zipf":
            keys = W.keys_for("zipf", n, B, seed)
        elif src == "delta":
            keys = W.delta_keys(n, B, param, seed)
        else:
            keys = (W.uniform_keys(n, seed) % np.uint32(B + 4)).astype(np.uint32)
        # values = original index: the output values are the stable permutation
        vals = np.arange(n, dtype=np.uint32)
        wk, wv, wo = oracle.multisplit(keys, B, kind, param, vals)
        nk, nv, no = multisplit_np(keys, B, kind, param, vals)
        check(wk, nk, name + " keys")
        check(wo, no, name + " offsets")
        check(wv, nv, name + " permutation")
        check(keys[wv], wk, name + " keys = keys[permutation]")
        out[f"{name}__meta"] = np.array([n, B, kind, param, int(kv)], dtype=np.uint32)
        out[f"{name}__keys"] = keys
        out[f"{name}__perm"] = wv.astype(np.uint16)
        out[f"{name}__offsets"] = wo

    # radix sort, 7- and 8-bit digits, key-value with value = index
    for bits in (7, 8):
        keys = W.uniform_keys(2003, 9100 + bits)
        keys[::7] = keys[1]
        vals = np.arange(keys.size, dtype=np.uint32)
        wk, wv = oracle.sort(keys, bits, vals)
        order = np.argsort(keys, kind="stable")
        check(wk, keys[order], f"sort{bits} keys")
        check(wv, vals[order], f"sort{bits} values")
        out[f"sort{bits}__keys"] = keys
        out[f"sort{bits}__perm"] = wv.astype(np.uint16)

    # 256-bin histogram of the low 8 bits; even float histogram; compaction
    keys = W.uniform_keys(4099, 9200)
    h = oracle.histogram(keys, 256, oracle.DIGIT, 0)
    check(h, np.bincount(keys & np.uint32(255), minlength=256).astype(np.uint32), "histogram")
    out["hist256__keys"] = keys
    out["hist256__hist"] = h
    x = (W.rng(9201).random(4099, dtype=np.float32) * np.float32(256.0)).astype(np.float32)
    inv = np.float32(32) / np.float32(256)
    he = oracle.histogram_even_f32(x, 32, 0.0, float(inv))
    bins = ((x - np.float32(0)) * inv).astype(np.float32).astype(np.int64)
    ok = (bins >= 0) & (bins < 32)
    check(he, np.bincount(bins[ok], minlength=32).astype(np.int32), "histogram even")
    out["histeven32__samples"] = x
    out["histeven32__hist"] = he
    keys = W.permutation_keys(6145, 9202)
    c, wk, _ = oracle.compaction(keys, oracle.LOWBITS, 1)
    keep = (keys & 1) == 0
    check(wk, np.concatenate([keys[keep], keys[~keep]]), "compaction")
    if c != int(keep.sum()):
        raise SystemExit("compaction count disagrees; nothing written")
    out["compact__keys"] = keys
    out["compact__out_keys"] = wk.astype(np.uint16)
    out["compact__count"] = np.array([c], dtype=np.uint32)

    # 2^20 keys, 8 buckets, key-only: seed + digest of the expected output
    n = 1 << 20
    keys = W.uniform_keys(n, 20)
    wk, _, wo = oracle.multisplit(keys, 8, oracle.RANGE, 0, None)
    nk, _, no = multisplit_np(keys, 8, oracle.RANGE, 0, None)
    check(wk, nk, "2^20 keys")
    check(wo, no, "2^20 offsets")
    out["big8__meta"] = np.array([n, 8, 20], dtype=np.uint32)
    out["big8__sha256_out_keys"] = np.frombuffer(hashlib.sha256(wk.tobytes()).digest(), np.uint8)
    out["big8__offsets"] = wo

    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
