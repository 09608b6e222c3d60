"""C-ABI library checks that need no GPU: the library loads, exports exactly the
symbols include/gpumultisplit.h declares, and its host-side logic (bucket-id
arithmetic, tile policy, argument validation before any HIP call) is right."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpumultisplit.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gms_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_reports_abi_version():
    L = gms.load_library()
    assert L.gms_abi_version() == 1


def test_exports_match_header():
    declared = header_functions()
    out = subprocess.run(["nm", "-D", "--defined-only", gms.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines() if " T gms_" in ln})
    assert declared == exported
    assert sorted(gms.EXPORTED_SYMBOLS) == declared
    L = gms.load_library()
    for name in declared:
        assert getattr(L, name) is not None


def test_status_strings():
    L = gms.load_library()
    for code in range(9):
        assert L.gms_status_string(code)
    assert L.gms_status_string(0).decode() == "success"


@pytest.mark.parametrize("delta", [1, 2, 3, 7, 10, 1000, 32768, 1048576, (1 << 25) // 32,
                                   0x7FFFFFFF, 0x80000000, 0x80000001, 0xFFFFFFFE, 0xFFFFFFFF,
                                   12345679, 641, 6700417])
def test_delta_magic_division_is_exact(delta):
    # The kernels divide by multiplying with a precomputed reciprocal; the host
    # entry point runs the same code.  Check against exact integer division on
    # keys around every multiple of delta and at the ends of the key range.
    g = W.rng(delta % 1000 + 1)
    keys = [0, 1, 2, 0xFFFFFFFF, 0xFFFFFFFE, 0x80000000, 0x7FFFFFFF]
    for m in list(range(0, 64)) + list(g.integers(0, (1 << 32) // delta + 1, 64)):
        for o in (-1, 0, 1):
            k = int(m) * delta + o
            if 0 <= k < (1 << 32):
                keys.append(k)
    keys = np.array(keys + list(W.uniform_keys(20000, seed=delta % 977)), dtype=np.uint64).astype(np.uint32)
    ids = gms.bucket_ids_host(keys, 256, gms.delta_func(delta))
    want = np.minimum(keys.astype(np.uint64) // delta, 255).astype(np.uint32)
    np.testing.assert_array_equal(ids, want)


@pytest.mark.parametrize("B", [1, 2, 3, 7, 8, 100, 255, 256])
def test_range_and_other_ids_match_oracle(B):
    keys = np.concatenate([W.uniform_keys(5000, seed=B),
                           np.array([0, 1, 0xFFFFFFFF, 0x80000000], dtype=np.uint32)])
    np.testing.assert_array_equal(gms.bucket_ids_host(keys, B, gms.range_bucket(B)),
                                  oracle.bucket_ids(keys, B, oracle.RANGE))
    np.testing.assert_array_equal(gms.bucket_ids_host(keys, B, gms.identity_bucket()),
                                  oracle.bucket_ids(keys, B, oracle.IDENTITY))
    if B & (B - 1) == 0 and B > 1:
        bits = B.bit_length() - 1
        digit = gms._digit_functor(bits)
        for shift in (0, 5, 32 - bits):
            np.testing.assert_array_equal(gms.bucket_ids_host(keys, B, digit(shift)),
                                          oracle.bucket_ids(keys, B, oracle.DIGIT, shift))
        np.testing.assert_array_equal(gms.bucket_ids_host(keys, B, gms.test_bucket(bits)),
                                      oracle.bucket_ids(keys, B, oracle.LOWBITS, bits))


def test_digit_functors_match_oracle():
    keys = W.uniform_keys(4000, seed=44)
    for bits, cls in ((2, gms.ms_bucket_sort_2bit), (4, gms.ms_bucket_sort_4bit),
                      (5, gms.ms_bucket_sort_5bit), (6, gms.ms_bucket_sort_6bit),
                      (7, gms.ms_bucket_sort_7bit), (8, gms.ms_bucket_sort_8bit)):
        for shift in (0, 3, 32 - bits):
            f = cls(shift)
            got = gms.bucket_ids_host(keys, 1 << bits, f)
            np.testing.assert_array_equal(got, oracle.bucket_ids(keys, 1 << bits, oracle.DIGIT, shift))
            assert all(int(got[i]) == f(int(keys[i])) for i in range(50))


def test_descriptor_validation_on_host():
    keys = np.zeros(4, dtype=np.uint32)
    with pytest.raises(gms.GmsError) as e:
        gms.bucket_ids_host(keys, 0, gms.range_bucket(0))
    assert e.value.status == 2
    with pytest.raises(gms.GmsError):
        gms.bucket_ids_host(keys, (1 << 20) + 1, gms.range_bucket((1 << 20) + 1))
    big = W.uniform_keys(1000, seed=257)                                # > 256 buckets: host check too
    for B in (257, 4096, 1 << 20):
        np.testing.assert_array_equal(gms.bucket_ids_host(big, B, gms.range_bucket(B)),
                                      oracle.bucket_ids(big, B, oracle.RANGE))
    with pytest.raises(gms.GmsError):
        gms.bucket_ids_host(keys, 6, gms.ms_bucket_sort_4bit(0))      # 16 digit values > 6 buckets
    np.testing.assert_array_equal(                                    # 4 digit values fit 6 buckets
        gms.bucket_ids_host(np.arange(8, dtype=np.uint32), 6, gms.ms_bucket_sort_2bit(1)),
        np.array([0, 0, 1, 1, 2, 2, 3, 3], dtype=np.uint32))
    with pytest.raises(gms.GmsError):
        gms.bucket_ids_host(keys, 2, gms.test_bucket(2))              # 4 ids > 2 buckets
    with pytest.raises(ValueError):
        gms.delta_func(0)


def test_tile_policy():
    for B in (2, 8, 32, 64, 256):
        t_ko = gms.subproblem_size_bms_key_only(B)
        t_kv = gms.subproblem_size_bms_key_value(B)
        assert t_ko % 256 == 0 and t_kv % 256 == 0 and 1024 <= t_kv <= 32768
    assert gms.lib().gms_subproblem_size(0, 1) == 0
    assert gms.lib().gms_subproblem_size(257, 1) == 0


def test_context_creation_checks_bucket_count():
    with pytest.raises(gms.GmsError) as e:
        gms.multisplit_context(0)
    assert e.value.status == 2
    with pytest.raises(gms.GmsError):
        gms.multisplit_context((1 << 20) + 1)
    assert gms.multisplit_context(300).num_buckets == 300             # reduced-bit sort path
    ctx = gms.multisplit_context(256)       # host-only object, no HIP call yet
    assert ctx.num_buckets == 256


def test_library_has_no_undefined_library_symbols():
    """Every gms:: symbol the library references is defined in it (a shared
    library links even when one is missing; loading then fails)."""
    r = subprocess.run(["nm", "-D", "-C", "--undefined-only", gms.LIB_PATH], capture_output=True,
                       text=True, check=True)
    missing = [ln for ln in r.stdout.splitlines() if "gms::" in ln or " gms_" in ln]
    assert not missing, missing


# 64-bit identifiers: the host entry point runs the kernels' AnyF64 code.
def _py_bucket64(kind, B, param, width, k):
    if kind == gms.BUCKET_RANGE:
        b = (k * B) >> 64
    elif kind == gms.BUCKET_DELTA:
        b = k // param
    elif kind == gms.BUCKET_DIGIT:
        b = (k >> param) & (((1 << width) - 1) if width else B - 1)
    elif kind == gms.BUCKET_IDENTITY:
        b = k
    else:
        b = k & ((1 << param) - 1)
    return min(b, B - 1)


def _ids64(B, kind, param, width, keys):
    L = gms.load_library()
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    ids = np.zeros(max(k.size, 1), np.uint32)
    st = L.gms_bucket_ids64_host(B, gms.gms_bucket_desc64(kind, width, param), k.ctypes.data,
                                 ids.ctypes.data, k.size)
    assert st == 0
    return ids[:k.size]


@pytest.mark.parametrize("delta", [1, 2, 3, 7, 1000, (1 << 32) - 1, 1 << 32, (1 << 32) + 1,
                                   (1 << 56) + 12345, (1 << 63) - 1, 1 << 63, (1 << 63) + 1,
                                   (1 << 64) - 2, (1 << 64) - 1, 6700417 * 641, 0x9E3779B97F4A7C15])
def test_delta64_magic_division_is_exact(delta):
    g = W.rng(delta % 997 + 3)
    keys = [0, 1, (1 << 64) - 1, (1 << 64) - 2, 1 << 63, (1 << 63) - 1]
    top = ((1 << 64) - 1) // delta
    for m in list(range(0, 40)) + [int(x) for x in g.integers(0, top + 1, 40, dtype=np.uint64)] + [top]:
        for o in (-1, 0, 1):
            k = m * delta + o
            if 0 <= k < (1 << 64):
                keys.append(k)
    keys += [int(x) for x in g.integers(0, 1 << 64, 200, dtype=np.uint64, endpoint=False)]
    got = _ids64(256, gms.BUCKET_DELTA, delta, 0, keys)
    want = [min(k // delta, 255) for k in keys]
    np.testing.assert_array_equal(got, np.array(want, dtype=np.uint32))
    # unclamped quotient check through a 1-bucket-per-quotient view: B = 256 and
    # keys whose quotient is below 256
    small = [q * delta + r for q in range(0, 256, 17) for r in (0, delta - 1) if q * delta + r < (1 << 64)]
    np.testing.assert_array_equal(_ids64(256, gms.BUCKET_DELTA, delta, 0, small),
                                  np.array([min(k // delta, 255) for k in small], dtype=np.uint32))


@pytest.mark.parametrize("B", [2, 3, 100, 255, 256])
def test_range64_boundaries(B):
    bnd = [(j << 64) // B for j in range(B)] + [-(-(j << 64) // B) for j in range(1, B)]
    keys = sorted({v + d for v in bnd for d in (-1, 0, 1) if 0 <= v + d < (1 << 64)} | {(1 << 64) - 1})
    np.testing.assert_array_equal(_ids64(B, gms.BUCKET_RANGE, 0, 0, keys),
                                  np.array([_py_bucket64(gms.BUCKET_RANGE, B, 0, 0, k) for k in keys],
                                           dtype=np.uint32))


@pytest.mark.parametrize("kind,B,param,width", [
    (gms.BUCKET_DIGIT, 256, 56, 8), (gms.BUCKET_DIGIT, 16, 60, 4), (gms.BUCKET_DIGIT, 64, 3, 0),
    (gms.BUCKET_DIGIT, 2, 63, 1), (gms.BUCKET_IDENTITY, 200, 0, 0), (gms.BUCKET_LOWBITS, 32, 5, 0),
])
def test_other_identifiers64(kind, B, param, width):
    g = W.rng(B + param)
    keys = [0, 1, (1 << 64) - 1, 1 << 63, 199, 200, 255, 256] + \
        [int(x) for x in g.integers(0, 1 << 64, 300, dtype=np.uint64, endpoint=False)]
    np.testing.assert_array_equal(_ids64(B, kind, param, width, keys),
                                  np.array([_py_bucket64(kind, B, param, width, k) for k in keys],
                                           dtype=np.uint32))


def test_wide_entry_points_validate_before_any_hip_call():
    L = gms.load_library()
    h = ctypes.c_void_p()
    assert L.gms_multisplit_create(256, ctypes.byref(h)) == 0
    d = gms.gms_bucket_desc64(gms.BUCKET_RANGE, 0, 0)
    # not allocated, bad widths, shift past the key, delta 0
    assert L.gms_multisplit_wide(h, 8, None, 16, None, 1, 8, 0, d, None, None, None) == 4
    assert L.gms_multisplit_wide(h, 8, None, 16, None, 1, 6, 0, d, None, None, None) == 1
    assert L.gms_multisplit_wide(h, 8, None, 16, None, 1, 8, 2, d, None, None, None) == 1
    assert L.gms_multisplit_wide(h, 8, 24, 16, None, 1, 8, 0, d, None, None, None) == 1
    assert L.gms_multisplit_destroy(h) == 0
    assert L.gms_bucket_ids64_host(256, gms.gms_bucket_desc64(gms.BUCKET_DIGIT, 8, 64), None,
                                   None, 0) == 1
    assert L.gms_bucket_ids64_host(256, gms.gms_bucket_desc64(gms.BUCKET_DELTA, 0, 0), None,
                                   None, 0) == 1
    assert L.gms_bucket_ids64_host(257, d, None, None, 0) == 2
    s = ctypes.c_void_p()
    assert L.gms_sort_create(ctypes.byref(s)) == 0
    assert L.gms_sort_wide(s, 8, 16, None, None, 1, 8, 0, None) == 4
    assert L.gms_sort_allocate_wide(s, 10, 3, 8, 0) == 1  # digit width outside 4..8
    assert L.gms_sort_allocate_wide(s, 10, 8, 8, 2) == 1
    assert L.gms_sort_destroy(s) == 0


def test_strided_and_expanded_buffers_rejected_before_any_launch():
    """Kernels take a buffer's address and read n consecutive elements, so a
    strided view or an expanded tensor is refused on the host, before the
    library is called (this runs without a GPU)."""
    import torch
    base = torch.zeros(64, dtype=torch.int32)
    ctx = gms.multisplit_context(4)
    for bad in (base[::2], base.view(8, 8)[:, 0], torch.zeros(1, dtype=torch.int32).expand(32)):
        assert not bad.is_contiguous()
        with pytest.raises(ValueError, match="contiguous"):
            gms.multisplit_key_only(bad, torch.zeros(32, dtype=torch.int32), bad.numel(), ctx,
                                    gms.range_bucket(4))
        with pytest.raises(ValueError, match="contiguous"):
            gms._ptr(bad)
    offs = torch.zeros(8, dtype=torch.int32)[::2]
    with pytest.raises(ValueError, match="contiguous"):
        gms._offsets_buffer(offs, 4)
