"""CPU oracle for the multisplit hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or, for
``cpu_baseline``, as the timed CPU restatement).  The product package
``gpumultisplit_amd`` never imports it.

numpy wrappers over ``oracle/liboracle.so`` (built from ``oracle/oracle.c`` by
``oracle/Makefile``).  The C file is a clean-room restatement of the
reference's CPU multisplit (src/cpu_functions.h:155-288), its sort digit
schedules (src/api/multisplit_sort.cuh:320-559), histogram semantics
(src/kernels/histogram/multisplit2_histograms.cuh:36-392) and compaction
return value (src/kernels/compaction/multisplit2_compaction.cuh:525-573).

Pinning: every function is checked in tests/test_oracle.py against an
independent computation (numpy stable sorts, bincount, Python integer
arithmetic, brute-force closed forms on tiny inputs).  No function is left
"parity unpinned".
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

RANGE, DELTA, DIGIT, IDENTITY, LOWBITS = 0, 1, 2, 3, 4

_lib = None


def build() -> None:
    """Compile oracle/oracle.c into oracle/liboracle.so."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u32p = ctypes.c_void_p
        L.ms_oracle_bucket_ids.argtypes = [ctypes.c_uint32] * 3 + [u32p, ctypes.c_uint64, u32p]
        L.ms_oracle_bucket_ids.restype = None
        L.ms_oracle_multisplit.argtypes = [ctypes.c_uint32] * 3 + [u32p] * 4 + [ctypes.c_uint64, u32p]
        L.ms_oracle_multisplit.restype = ctypes.c_int
        L.ms_oracle_sort_schedule.argtypes = [ctypes.c_uint32, u32p, u32p]
        L.ms_oracle_sort_schedule.restype = ctypes.c_int
        L.ms_oracle_sort.argtypes = [u32p] * 4 + [ctypes.c_uint64, ctypes.c_uint32]
        L.ms_oracle_sort.restype = ctypes.c_int
        L.ms_oracle_histogram.argtypes = [ctypes.c_uint32] * 3 + [u32p, ctypes.c_uint64, u32p]
        L.ms_oracle_histogram.restype = None
        L.ms_oracle_histogram_even_f32.argtypes = [u32p, ctypes.c_uint64, u32p, ctypes.c_uint32,
                                                   ctypes.c_float, ctypes.c_float]
        L.ms_oracle_histogram_even_f32.restype = None
        for name in ("ms_oracle_histogram_range_f32", "ms_oracle_histogram_range_u32"):
            fn = getattr(L, name)
            fn.argtypes = [u32p, ctypes.c_uint64, u32p, u32p, ctypes.c_uint32]
            fn.restype = None
        u64 = ctypes.c_uint64
        L.ms_oracle_bucket_ids64.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u64, ctypes.c_uint32,
                                             u32p, u64, u32p]
        L.ms_oracle_bucket_ids64.restype = None
        L.ms_oracle_multisplit_wide.argtypes = ([ctypes.c_uint32, ctypes.c_uint32, u64] +
                                                [ctypes.c_uint32] * 3 + [u32p] * 4 + [u64, u32p])
        L.ms_oracle_multisplit_wide.restype = ctypes.c_int
        L.ms_oracle_sort_wide.argtypes = [u32p] * 4 + [u64, ctypes.c_uint32, ctypes.c_uint32]
        L.ms_oracle_sort_wide.restype = ctypes.c_int
        L.ms_oracle_compaction.argtypes = [ctypes.c_uint32] * 2 + [u32p] * 4 + [ctypes.c_uint64]
        L.ms_oracle_compaction.restype = ctypes.c_uint64
        L.ms_oracle_tile_histograms.argtypes = [ctypes.c_uint32] * 3 + [u32p, ctypes.c_uint64,
                                                                        ctypes.c_uint32, u32p]
        L.ms_oracle_tile_histograms.restype = None
        L.ms_baseline_stable_sort.argtypes = [u32p] * 4 + [ctypes.c_uint64]
        L.ms_baseline_stable_sort.restype = None
        L.ms_baseline_histogram.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                            u32p]
        L.ms_baseline_histogram.restype = None
        L.ms_baseline_stable_partition.argtypes = [u32p] * 4 + [ctypes.c_uint64, ctypes.c_uint64]
        L.ms_baseline_stable_partition.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a):
    return a.ctypes.data if a is not None else None


def bucket_ids(keys, num_buckets: int, kind: int = RANGE, param: int = 0) -> np.ndarray:
    keys = _u32(keys)
    ids = np.empty_like(keys)
    lib().ms_oracle_bucket_ids(kind, num_buckets, param, _ptr(keys), keys.size, _ptr(ids))
    return ids


def multisplit(keys, num_buckets: int, kind: int = RANGE, param: int = 0, values=None):
    """Stable multisplit.  Returns (keys_out, values_out or None, offsets[B])."""
    keys = _u32(keys)
    vals = _u32(values) if values is not None else None
    kout = np.empty_like(keys)
    vout = np.empty_like(vals) if vals is not None else None
    offs = np.zeros(num_buckets, dtype=np.uint32)
    rc = lib().ms_oracle_multisplit(kind, num_buckets, param, _ptr(keys), _ptr(vals),
                                    _ptr(kout), _ptr(vout), keys.size, _ptr(offs))
    if rc != 0:
        raise MemoryError("oracle multisplit failed")
    return kout, vout, offs


def sort_schedule(digit_bits: int):
    s = np.zeros(8, dtype=np.uint32)
    b = np.zeros(8, dtype=np.uint32)
    p = lib().ms_oracle_sort_schedule(digit_bits, _ptr(s), _ptr(b))
    return [(int(s[i]), int(b[i])) for i in range(p)]


def sort(keys, digit_bits: int = 8, values=None):
    keys = _u32(keys)
    vals = _u32(values) if values is not None else None
    kout = np.empty_like(keys)
    vout = np.empty_like(vals) if vals is not None else None
    rc = lib().ms_oracle_sort(_ptr(keys), _ptr(vals), _ptr(kout), _ptr(vout), keys.size, digit_bits)
    if rc != 0:
        raise ValueError(f"oracle sort failed ({rc})")
    return kout, vout


def histogram(keys, num_buckets: int, kind: int = RANGE, param: int = 0) -> np.ndarray:
    keys = _u32(keys)
    h = np.zeros(num_buckets, dtype=np.uint32)
    lib().ms_oracle_histogram(kind, num_buckets, param, _ptr(keys), keys.size, _ptr(h))
    return h


def histogram_even_f32(samples, num_bins: int, lower: float, inv_delta: float) -> np.ndarray:
    x = np.ascontiguousarray(samples, dtype=np.float32)
    h = np.zeros(num_bins, dtype=np.int32)
    lib().ms_oracle_histogram_even_f32(_ptr(x), x.size, _ptr(h), num_bins,
                                       ctypes.c_float(lower), ctypes.c_float(inv_delta))
    return h


def histogram_range(samples, levels) -> np.ndarray:
    """Counts per bin [levels[i], levels[i+1]) of float32 or uint32 samples."""
    if np.asarray(samples).dtype == np.float32:
        x = np.ascontiguousarray(samples, dtype=np.float32)
        lv = np.ascontiguousarray(levels, dtype=np.float32)
        fn = lib().ms_oracle_histogram_range_f32
    else:
        x = _u32(samples)
        lv = _u32(levels)
        fn = lib().ms_oracle_histogram_range_u32
    h = np.zeros(max(lv.size - 1, 0), dtype=np.int32)
    fn(_ptr(x), x.size, _ptr(h), _ptr(lv), lv.size)
    return h


def compaction(keys, kind: int = LOWBITS, param: int = 1, values=None):
    keys = _u32(keys)
    vals = _u32(values) if values is not None else None
    kout = np.empty_like(keys)
    vout = np.empty_like(vals) if vals is not None else None
    c = lib().ms_oracle_compaction(kind, param, _ptr(keys), _ptr(vals), _ptr(kout), _ptr(vout), keys.size)
    return int(c), kout, vout


def tile_histograms(keys, num_buckets: int, tile: int, kind: int = RANGE, param: int = 0) -> np.ndarray:
    keys = _u32(keys)
    nt = (keys.size + tile - 1) // tile
    H = np.zeros(num_buckets * max(nt, 1), dtype=np.uint32)
    lib().ms_oracle_tile_histograms(kind, num_buckets, param, _ptr(keys), keys.size, tile, _ptr(H))
    return H[: num_buckets * nt]


def _words(a):
    """(uint32 word view, words per element) of a uint32/uint64 array."""
    a = np.ascontiguousarray(a)
    if a.dtype.itemsize == 8:
        return np.ascontiguousarray(a.astype(np.uint64)).view(np.uint32), 2
    return _u32(a), 1


def bucket_ids64(keys, num_buckets: int, kind: int = RANGE, param: int = 0, width: int = 0) -> np.ndarray:
    """Bucket ids of 64-bit keys (the 64-bit forms of the identifiers)."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    ids = np.zeros(max(k.size, 1), np.uint32)
    lib().ms_oracle_bucket_ids64(kind, num_buckets, int(param), width, _ptr(k.view(np.uint32)),
                                 k.size, _ptr(ids))
    return ids[:k.size]


def multisplit_wide(keys, num_buckets: int, kind: int = RANGE, param: int = 0, values=None,
                    width: int = 0):
    """Stable multisplit of uint64 (or uint32) keys carrying uint32/uint64 values.
    Returns (keys_out, values_out or None, bucket starts)."""
    kw, nkw = _words(keys)
    n = kw.size // nkw
    vw, nvw = (None, 0) if values is None else _words(values)
    ko = np.zeros(max(kw.size, 1), np.uint32)
    vo = None if vw is None else np.zeros(max(vw.size, 1), np.uint32)
    offs = np.zeros(num_buckets, np.uint32)
    rc = lib().ms_oracle_multisplit_wide(kind, num_buckets, int(param), width, nkw, nvw, _ptr(kw),
                                         _ptr(vw), _ptr(ko), _ptr(vo), n, _ptr(offs))
    assert rc == 0
    kt = np.uint64 if nkw == 2 else np.uint32
    kout = ko[:kw.size].view(kt)
    vout = None if vo is None else vo[:vw.size].view(np.uint64 if nvw == 2 else np.uint32)
    return kout, vout, offs


def sort_wide(keys, digit_bits: int = 8, values=None):
    """Stable LSD sort of uint64 keys (values of 32 or 64 bits follow)."""
    kw, _ = _words(np.asarray(keys, dtype=np.uint64))
    n = kw.size // 2
    vw, nvw = (None, 0) if values is None else _words(values)
    ko = np.zeros(max(kw.size, 2), np.uint32)
    vo = None if vw is None else np.zeros(max(vw.size, 1), np.uint32)
    rc = lib().ms_oracle_sort_wide(_ptr(kw), _ptr(vw), _ptr(ko), _ptr(vo), n, nvw, digit_bits)
    assert rc == 0
    vout = None if vo is None else vo[:vw.size].view(np.uint64 if nvw == 2 else np.uint32)
    return ko[:kw.size].view(np.uint64), vout


# ---------------------------------------------------------------------------
# Standard-library CPU baselines (oracle/baseline.cpp): what bench.py times on
# the host for the sort, histogram and compaction rows.
# ---------------------------------------------------------------------------
def baseline_stable_sort(keys, values=None):
    """std::stable_sort of (key, value) pairs by key (keys alone if no values)."""
    keys = _u32(keys)
    ko = np.empty_like(keys)
    vals = None if values is None else _u32(values)
    vo = None if vals is None else np.empty_like(vals)
    lib().ms_baseline_stable_sort(_ptr(keys), _ptr(vals), _ptr(ko), _ptr(vo), keys.size)
    return ko, vo


def baseline_histogram(keys, bins: int = 256, shift: int = 0) -> np.ndarray:
    """Counting loop over (key >> shift) & (bins - 1)."""
    keys = _u32(keys)
    h = np.zeros(bins, dtype=np.uint32)
    lib().ms_baseline_histogram(_ptr(keys), keys.size, shift, bins, _ptr(h))
    return h


def baseline_stable_partition(keys, threshold: int, values=None):
    """std::stable_partition by key < threshold.  Returns (count, keys, values)."""
    keys = _u32(keys)
    ko = np.empty_like(keys)
    vals = None if values is None else _u32(values)
    vo = None if vals is None else np.empty_like(vals)
    c = lib().ms_baseline_stable_partition(_ptr(keys), _ptr(vals), _ptr(ko), _ptr(vo), keys.size,
                                           threshold)
    return int(c), ko, vo
