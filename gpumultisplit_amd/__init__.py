"""gpumultisplit_amd -- MI355X (gfx950) multisplit, radix sort, histogram and
compaction behind the reference's API names.

The compute path is ``libgpumultisplit.so`` (hand-written HIP kernels, see
``csrc/``), reached through its C ABI (``include/gpumultisplit.h``) with
ctypes.  This module mirrors the reference's C++ host API
(src/api/multisplit.cuh, src/api/multisplit_sort.cuh,
src/api/multisplit_sort_pairs.cuh,
src/kernels/compaction/multisplit2_compaction.cuh) with the same names and
argument order.  Device buffers are torch tensors on the GPU (any 32-bit
dtype; the bits are moved) or raw integer device pointers.

Differences from the reference, all deliberate:
* errors raise ``GmsError`` instead of printing or silently doing nothing;
* one allocation serves every call with n <= the allocated n;
* ``multisplit_offset`` (host array of B bucket starts) is filled with one
  copy instead of B-1 blocking copies.

There is no CPU fallback: when the library is missing every call raises.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

from . import workloads  # noqa: F401  (seeded synthetic inputs)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpumultisplit.so")

GMS_SUCCESS = 0
STATUS_NAMES = {
    0: "GMS_SUCCESS", 1: "GMS_ERR_INVALID_ARG", 2: "GMS_ERR_UNSUPPORTED_BUCKETS",
    3: "GMS_ERR_CAPACITY", 4: "GMS_ERR_NOT_ALLOCATED", 5: "GMS_ERR_ALLOC", 6: "GMS_ERR_HIP",
    7: "GMS_ERR_ALIAS", 8: "GMS_ERR_NO_DEVICE",
}

BUCKET_RANGE, BUCKET_DELTA, BUCKET_DIGIT, BUCKET_IDENTITY, BUCKET_LOWBITS = 0, 1, 2, 3, 4
VARIANT_AUTO, VARIANT_BMS, VARIANT_WMS, VARIANT_DMS = 0, 1, 2, 3


class GmsError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = STATUS_NAMES.get(status, str(status))
        if _lib is not None:
            msg += ": " + _lib.gms_status_string(status).decode()
        super().__init__(f"{what}: {msg}" if what else msg)


class gms_bucket_desc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("param", ctypes.c_uint32), ("width", ctypes.c_uint32)]


class gms_bucket_desc64(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("width", ctypes.c_uint32), ("param", ctypes.c_uint64)]


class gms_stage_plan(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint32) for f in (
        "num_buckets", "log_b", "key_words", "value_words", "n", "num_tiles", "large", "vec",
        "done")] + [
        ("d_matrix", ctypes.c_void_p), ("d_partials", ctypes.c_void_p),
        ("d_staging", ctypes.c_void_p), ("scan_words", ctypes.c_uint64)]


class gms_batch_item(ctypes.Structure):
    _fields_ = [("d_key_in", ctypes.c_void_p), ("d_value_in", ctypes.c_void_p),
                ("d_key_out", ctypes.c_void_p), ("d_value_out", ctypes.c_void_p),
                ("n", ctypes.c_uint32), ("num_buckets", ctypes.c_uint32),
                ("desc", gms_bucket_desc), ("d_offsets", ctypes.c_void_p)]


class gms_batch_sort_item(ctypes.Structure):
    _fields_ = [("d_key_in", ctypes.c_void_p), ("d_value_in", ctypes.c_void_p),
                ("d_key_out", ctypes.c_void_p), ("d_value_out", ctypes.c_void_p),
                ("n", ctypes.c_uint32), ("digit_bits", ctypes.c_uint32),
                ("key_bytes", ctypes.c_uint32), ("value_bytes", ctypes.c_uint32)]


class gms_batch_wide_item(ctypes.Structure):
    _fields_ = [("d_key_in", ctypes.c_void_p), ("d_value_in", ctypes.c_void_p),
                ("d_key_out", ctypes.c_void_p), ("d_value_out", ctypes.c_void_p),
                ("n", ctypes.c_uint32), ("num_buckets", ctypes.c_uint32),
                ("key_bytes", ctypes.c_uint32), ("value_bytes", ctypes.c_uint32),
                ("desc", gms_bucket_desc64), ("d_offsets", ctypes.c_void_p)]


class gms_batch_tuning(ctypes.Structure):
    _fields_ = [("items_per_thread", ctypes.c_uint32), ("persistent", ctypes.c_uint32)]


class gms_request(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint32), ("d_key_in", ctypes.c_void_p),
                ("d_value_in", ctypes.c_void_p), ("d_key_out", ctypes.c_void_p),
                ("d_value_out", ctypes.c_void_p), ("n", ctypes.c_uint32),
                ("num_buckets", ctypes.c_uint32), ("desc", gms_bucket_desc),
                ("digit_bits", ctypes.c_uint32), ("d_offsets", ctypes.c_void_p),
                ("key_bytes", ctypes.c_uint32), ("value_bytes", ctypes.c_uint32),
                ("desc64", gms_bucket_desc64)]


class gms_server_config(ctypes.Structure):
    _fields_ = [("scratch_budget_bytes", ctypes.c_uint64), ("max_inflight_steps", ctypes.c_uint32),
                ("max_step_keys", ctypes.c_uint32), ("items_per_thread", ctypes.c_uint32)]


class gms_server_stats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("step_items", ctypes.c_uint64), ("keys_processed", ctypes.c_uint64),
                ("peak_scratch_bytes", ctypes.c_uint64), ("active", ctypes.c_uint32),
                ("pending", ctypes.c_uint32)]


_lib = None

# (name, restype, argtypes) of every exported symbol, in header order.
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I = ctypes.c_int
_PP = ctypes.POINTER(ctypes.c_void_p)
_SIGS = [
    ("gms_status_string", ctypes.c_char_p, [_I]),
    ("gms_abi_version", _U32, []),
    ("gms_bucket_ids_host", _I, [_U32, gms_bucket_desc, _P, _P, ctypes.c_size_t]),
    ("gms_bucket_ids64_host", _I, [_U32, gms_bucket_desc64, _P, _P, ctypes.c_size_t]),
    ("gms_multisplit_create", _I, [_U32, _PP]),
    ("gms_multisplit_allocate", _I, [_P, _U32, _I]),
    ("gms_multisplit_release", _I, [_P]),
    ("gms_multisplit_destroy", _I, [_P]),
    ("gms_multisplit_key_only", _I, [_P, _P, _P, _U32, gms_bucket_desc, _P, _P]),
    ("gms_multisplit_key_value", _I, [_P, _P, _P, _P, _P, _U32, gms_bucket_desc, _P, _P]),
    ("gms_multisplit", _I, [_P, _P, _P, _P, _P, _U32, gms_bucket_desc, _P, _P]),
    ("gms_multisplit_noalloc", _I, [_P, _P, _P, _P, _U32, _U32, gms_bucket_desc, _P, _P]),
    ("gms_multisplit_wide", _I, [_P, _P, _P, _P, _P, _U32, _U32, _U32, gms_bucket_desc64, _P, _P,
                                 _P]),
    ("gms_multisplit_set_variant", _I, [_P, _I]),
    ("gms_subproblem_size_wms", _U32, [_U32, _I, ctypes.POINTER(_U32)]),
    ("gms_subproblem_size", _U32, [_U32, _I]),
    ("gms_tile_keys", _U32, [_U32, _I, _U32]),
    ("gms_multisplit_set_profiling", _I, [_P, _I]),
    ("gms_multisplit_profile", _I, [_P, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_U32)]),
    ("gms_multisplit_prescan_matrix", _I, [_P, _P, _U32, gms_bucket_desc, _I, _P, _P]),
    ("gms_exclusive_scan_u32", _I, [_P, _P, _P, _U32, _P]),
    ("gms_multisplit_stage_plan", _I, [_P, _P, _P, _P, _P, _U32, _U32, _U32, _P,
                                       ctypes.POINTER(gms_stage_plan)]),
    ("gms_stage_scan", _I, [ctypes.POINTER(gms_stage_plan), _P]),
    ("gms_stage_offsets", _I, [ctypes.POINTER(gms_stage_plan), _P, _P]),
    ("gms_sort_create", _I, [_PP]),
    ("gms_sort_allocate", _I, [_P, _U32, _U32, _I]),
    ("gms_sort_release", _I, [_P]),
    ("gms_sort_destroy", _I, [_P]),
    ("gms_sort", _I, [_P, _P, _P, _P, _P, _U32, _U32, _P]),
    ("gms_sort_allocate_wide", _I, [_P, _U32, _U32, _U32, _U32]),
    ("gms_sort_wide", _I, [_P, _P, _P, _P, _P, _U32, _U32, _U32, _P]),
    ("gms_compaction_create", _I, [_PP]),
    ("gms_compaction_allocate", _I, [_P, _U32, _I]),
    ("gms_compaction_release", _I, [_P]),
    ("gms_compaction_destroy", _I, [_P]),
    ("gms_compaction", _I, [_P, _P, _P, _P, _P, _U32, gms_bucket_desc, _P, _P, _P]),
    ("gms_compaction_wide", _I, [_P, _P, _P, _P, _P, _U32, _U32, _U32, gms_bucket_desc64, _P, _P,
                                 _P]),
    ("gms_compaction_stage_plan", _I, [_P, _P, _P, _P, _P, _U32, _U32, _U32, _P,
                                       ctypes.POINTER(gms_stage_plan)]),
    ("gms_histogram_u32", _I, [_P, _U32, _P, _U32, gms_bucket_desc, _P]),
    ("gms_histogram_even_f32", _I, [_P, _U32, _P, _U32, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, _P]),
    ("gms_histogram_wide", _I, [_P, _U32, _U32, _P, _U32, gms_bucket_desc64, _P]),
    ("gms_histogram_range_f32", _I, [_P, _U32, _P, _U32, _P, _P]),
    ("gms_histogram_range_u32", _I, [_P, _U32, _P, _U32, _P, _P]),
    ("gms_batch_create", _I, [_PP]),
    ("gms_batch_destroy", _I, [_P]),
    ("gms_batch_plan", _I, [_P, ctypes.POINTER(gms_batch_item), _U32, _P]),
    ("gms_batch_plan_mixed", _I, [_P, ctypes.POINTER(gms_batch_item), _U32,
                                  ctypes.POINTER(gms_batch_sort_item), _U32, _P]),
    ("gms_batch_plan_wide", _I, [_P, ctypes.POINTER(gms_batch_item), _U32,
                                 ctypes.POINTER(gms_batch_wide_item), _U32,
                                 ctypes.POINTER(gms_batch_sort_item), _U32, _P]),
    ("gms_batch_run", _I, [_P, _P]),
    ("gms_batch_capture", _I, [_P]),
    ("gms_batch_launch_graph", _I, [_P, _P]),
    ("gms_batch_autotune", _I, [_P, _U32, _P, ctypes.POINTER(gms_batch_tuning),
                                ctypes.POINTER(ctypes.c_float)]),
    ("gms_batch_set_tuning", _I, [_P, gms_batch_tuning]),
    ("gms_batch_get_tuning", _I, [_P, ctypes.POINTER(gms_batch_tuning)]),
    ("gms_batch_stats", _I, [_P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(_U32),
                             ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(_U32)]),
    ("gms_server_create", _I, [ctypes.POINTER(gms_server_config), _PP]),
    ("gms_server_destroy", _I, [_P]),
    ("gms_server_submit", _I, [_P, ctypes.POINTER(gms_request), ctypes.POINTER(ctypes.c_uint64)]),
    ("gms_server_poll", _I, [_P, ctypes.c_uint64, ctypes.POINTER(_I),
                             ctypes.POINTER(ctypes.c_double)]),
    ("gms_server_wait", _I, [_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
    ("gms_server_get_stats", _I, [_P, ctypes.POINTER(gms_server_stats)]),
    ("gms_copy_u32", _I, [_P, _P, ctypes.c_size_t, _P]),
]
EXPORTED_SYMBOLS = [s[0] for s in _SIGS]


def build(verbose: bool = False) -> None:
    """Compile every HIP source for gfx950 into libgpumultisplit.so (in-tree)."""
    jobs = str(min(16, os.cpu_count() or 8))
    out = None if verbose else subprocess.DEVNULL
    subprocess.run(["make", "-j", jobs, "-C", os.path.join(_HERE, "csrc")], check=True, stdout=out)


def load_library():
    """Load libgpumultisplit.so.  Raises if it is missing: there is no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; run gpumultisplit_amd.build() "
                               "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def lib():
    return load_library()


def _check(status: int, what: str) -> None:
    if status != GMS_SUCCESS:
        raise GmsError(status, what)


def _require_contiguous(t):
    """The kernels read and write n consecutive elements from a buffer's
    address: a strided view would give them the wrong elements, and an
    expanded tensor (stride 0) has fewer elements of storage than it shows."""
    if hasattr(t, "is_contiguous") and not t.is_contiguous():
        raise ValueError("buffer must be contiguous (a strided or expanded view is not)")


def _ptr(x, sizes=(4,)):
    """Device pointer of a torch tensor (or an int address); None stays None.
    Tensors must have elements of one of `sizes` bytes."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        _require_contiguous(x)
        if hasattr(x, "is_cuda") and not x.is_cuda:
            raise ValueError("device buffer expected, got a host tensor")
        if x.element_size() not in sizes:
            raise ValueError(f"elements of {' or '.join(str(8 * b) for b in sizes)} bits expected")
        return x.data_ptr()
    raise TypeError(f"cannot take a device pointer of {type(x)!r}")


def _elem_bytes(*tensors):
    """Element size shared by the given tensors (4 for raw addresses and None)."""
    sizes = {t.element_size() for t in tensors if t is not None and hasattr(t, "element_size")}
    if len(sizes) > 1:
        raise ValueError("input and output buffers must have the same element size")
    return sizes.pop() if sizes else 4


def _is_wide(*tensors):
    return any(t is not None and hasattr(t, "element_size") and t.element_size() == 8
               for t in tensors)


def _stream(*tensors):
    for t in tensors:
        if t is not None and hasattr(t, "is_cuda") and t.is_cuda:
            import torch
            return torch.cuda.current_stream(t.device).cuda_stream
    return None


def _check_len(name, t, n):
    if t is not None and hasattr(t, "numel") and t.numel() < n:
        raise ValueError(f"{name} holds {t.numel()} elements, fewer than num_elements={n}")


# ---------------------------------------------------------------------------
# Bucket identifiers.  Same names as the reference's functors; each maps to a
# C-ABI descriptor and also evaluates on the host (for checks).
# ---------------------------------------------------------------------------
class _Bucket:
    kind = BUCKET_RANGE
    param = 0
    width = 0

    def desc(self) -> gms_bucket_desc:
        if self.param > 0xFFFFFFFF:
            raise ValueError("parameter needs 64 bits: use it with 64-bit keys")
        return gms_bucket_desc(self.kind, self.param, self.width)

    def desc64(self) -> gms_bucket_desc64:
        return gms_bucket_desc64(self.kind, self.width, self.param)


class range_bucket(_Bucket):
    """floor(key * B / 2^32) for B buckets (the sweep's range identifier)."""
    kind = BUCKET_RANGE

    def __init__(self, num_buckets: int):
        self.num_buckets = num_buckets

    def __call__(self, a: int) -> int:
        return (int(a) * self.num_buckets) >> 32


class delta_func(_Bucket):
    """key / delta (src/main/main_multisplit.cu:41-47)."""
    kind = BUCKET_DELTA

    def __init__(self, delta: int):
        if delta < 1:
            raise ValueError("delta must be >= 1")
        self.delta_ = self.param = int(delta)

    def __call__(self, a: int) -> int:
        return int(a) // self.delta_


class identity_bucket(_Bucket):
    """key itself (src/main/main_multisplit.cu:34-39)."""
    kind = BUCKET_IDENTITY

    def __call__(self, a: int) -> int:
        return int(a)


class test_bucket(_Bucket):
    """key & ((1 << flag) - 1) (src/main/main_compaction.cu:27-32)."""
    kind = BUCKET_LOWBITS

    def __init__(self, flag: int):
        self.flag = self.param = int(flag)

    def __call__(self, a: int) -> int:
        return int(a) & ((1 << self.flag) - 1)


def _digit_functor(bits: int):
    class _Digit(_Bucket):
        kind = BUCKET_DIGIT
        width = bits

        def __init__(self, x: int):
            self.x = self.param = int(x)

        def __call__(self, a: int) -> int:
            return (int(a) >> self.x) & ((1 << bits) - 1)

    _Digit.__name__ = f"ms_bucket_sort_{bits}bit"
    _Digit.__doc__ = f"(key >> x) & {(1 << bits) - 1} (src/api/multisplit_sort.cuh:63-98)."
    return _Digit


class digit_bucket(_Bucket):
    """(key >> shift) & (2^width - 1) for any width up to 20 bits (the
    ms_bucket_sort_Nbit functors fix width to 2..8)."""
    kind = BUCKET_DIGIT

    def __init__(self, shift: int, width: int):
        self.param, self.width = int(shift), int(width)

    def __call__(self, a: int) -> int:
        return (int(a) >> self.param) & ((1 << self.width) - 1)


ms_bucket_sort_2bit = _digit_functor(2)
ms_bucket_sort_4bit = _digit_functor(4)
ms_bucket_sort_5bit = _digit_functor(5)
ms_bucket_sort_6bit = _digit_functor(6)
ms_bucket_sort_7bit = _digit_functor(7)
ms_bucket_sort_8bit = _digit_functor(8)


def bucket_ids_host(keys, num_buckets: int, bucket_identifier):
    """Bucket ids computed by the library's own host copy of the kernel code."""
    import numpy as np
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    ids = np.empty_like(k)
    _check(lib().gms_bucket_ids_host(num_buckets, bucket_identifier.desc(), k.ctypes.data,
                                     ids.ctypes.data, k.size), "gms_bucket_ids_host")
    return ids


# ---------------------------------------------------------------------------
# Multisplit (src/api/multisplit.cuh)
# ---------------------------------------------------------------------------
def subproblem_size_bms_key_only(num_buckets: int) -> int:
    return int(lib().gms_subproblem_size(num_buckets, 0))


def subproblem_size_bms_key_value(num_buckets: int) -> int:
    return int(lib().gms_subproblem_size(num_buckets, 1))


def subproblem_size_wms_key_only(num_buckets: int):
    """(keys per wave subproblem, keys per workgroup) of the wave-level
    variant (src/api/wms_api.h:25-53); (0, 0) outside 2..64 buckets."""
    blk = _U32()
    sub = lib().gms_subproblem_size_wms(int(num_buckets), 0, ctypes.byref(blk))
    return int(sub), int(blk.value) if sub else 0


def subproblem_size_wms_key_value(num_buckets: int):
    """src/api/wms_api.h:55-83 (same shape as key-only here)."""
    blk = _U32()
    sub = lib().gms_subproblem_size_wms(int(num_buckets), 1, ctypes.byref(blk))
    return int(sub), int(blk.value) if sub else 0


def set_variant(context, variant: int) -> None:
    """Block-wide (VARIANT_BMS), wave-level (VARIANT_WMS, up to 64 buckets),
    wave-level with direct scatter (VARIANT_DMS, up to 64 buckets) or
    automatic (VARIANT_AUTO, which is BMS) kernels for the context's
    multisplits."""
    _check(lib().gms_multisplit_set_variant(context.handle, int(variant)), "set_variant")


def tile_keys(num_buckets: int, key_value: bool, num_elements: int) -> int:
    """Keys per tile of a call on num_elements keys."""
    return int(lib().gms_tile_keys(num_buckets, int(bool(key_value)), int(num_elements)))


class multisplit_context:
    """multisplit_context(n_buckets) (src/api/multisplit.cuh:29-47)."""

    def __init__(self, n_buckets: int):
        self.num_buckets = int(n_buckets)
        self.size_sub_prob = 0
        self.num_sub_prob = 0
        self.num_blocks = 0
        self.capacity = 0
        h = ctypes.c_void_p()
        _check(lib().gms_multisplit_create(self.num_buckets, ctypes.byref(h)), "multisplit_context")
        self._h = h

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if self._h and _lib is not None:
                _lib.gms_multisplit_destroy(self._h)
                self._h = None
        except Exception:
            pass


def _ms_allocate(num_elements: int, context: multisplit_context, kv: int) -> None:
    _check(lib().gms_multisplit_allocate(context.handle, int(num_elements), kv),
           "multisplit_allocate")
    context.capacity = int(num_elements)
    context.size_sub_prob = int(lib().gms_subproblem_size(context.num_buckets, kv))
    # the reference's subproblem sizes stop at 256 buckets (subproblem_size_bms):
    # 0 subproblems above
    context.num_sub_prob = context.num_blocks = (
        (int(num_elements) + context.size_sub_prob - 1) // context.size_sub_prob
        if context.size_sub_prob else 0)


def multisplit_allocate_key_only(num_elements: int, context: multisplit_context) -> None:
    """src/api/multisplit.cuh:50-61."""
    _ms_allocate(num_elements, context, 0)


def multisplit_allocate_key_value(num_elements: int, context: multisplit_context) -> None:
    """src/api/multisplit.cuh:63-74."""
    _ms_allocate(num_elements, context, 1)


def multisplit_release_memory(context: multisplit_context) -> None:
    """src/api/multisplit.cuh:76-79."""
    _check(lib().gms_multisplit_release(context.handle), "multisplit_release_memory")


def _offsets_buffer(multisplit_offset, num_buckets):
    if multisplit_offset is None:
        return None, None
    import numpy as np
    if isinstance(multisplit_offset, np.ndarray):
        if multisplit_offset.dtype != np.uint32 or not multisplit_offset.flags.c_contiguous:
            raise ValueError("multisplit_offset must be a contiguous uint32 array")
        if multisplit_offset.size < num_buckets:
            raise ValueError("multisplit_offset needs num_buckets entries")
        return multisplit_offset.ctypes.data, None
    if hasattr(multisplit_offset, "data_ptr") and not multisplit_offset.is_cuda:
        if multisplit_offset.numel() < num_buckets or multisplit_offset.element_size() != 4:
            raise ValueError("multisplit_offset needs num_buckets 32-bit entries")
        _require_contiguous(multisplit_offset)
        return multisplit_offset.data_ptr(), None
    raise TypeError("multisplit_offset must be a host numpy array or CPU tensor")


def _multisplit_wide(d_key_in, d_value_in, d_key_out, d_value_out, num_elements, context,
                     bucket_identifier, h_off, d_off, stream, what):
    kb = _elem_bytes(d_key_in, d_key_out)
    vb = 0 if d_value_in is None and d_value_out is None else _elem_bytes(d_value_in, d_value_out)
    _check(lib().gms_multisplit_wide(context.handle, _ptr(d_key_in, (4, 8)), _ptr(d_value_in, (4, 8)),
                                     _ptr(d_key_out, (4, 8)), _ptr(d_value_out, (4, 8)),
                                     int(num_elements), kb, vb, bucket_identifier.desc64(), d_off,
                                     h_off, stream), what)


def multisplit_key_only(d_key_in, d_key_out, num_elements: int, context: multisplit_context,
                        bucket_identifier, multisplit_offset=None) -> None:
    """Stable key-only multisplit (src/api/multisplit.cuh:81-197); 32- or
    64-bit keys (the reference's KeyT)."""
    _check_len("d_key_in", d_key_in, num_elements)
    _check_len("d_key_out", d_key_out, num_elements)
    h_off, _ = _offsets_buffer(multisplit_offset, context.num_buckets)
    if _is_wide(d_key_in, d_key_out):
        _multisplit_wide(d_key_in, None, d_key_out, None, num_elements, context, bucket_identifier,
                         h_off, None, _stream(d_key_in, d_key_out), "multisplit_key_only")
        return
    _check(lib().gms_multisplit_key_only(context.handle, _ptr(d_key_in), _ptr(d_key_out),
                                         int(num_elements), bucket_identifier.desc(), h_off,
                                         _stream(d_key_in, d_key_out)), "multisplit_key_only")


def multisplit_key_value(d_key_in, d_value_in, d_key_out, d_value_out, num_elements: int,
                         context: multisplit_context, bucket_identifier,
                         multisplit_offset=None) -> None:
    """Stable key-value multisplit (src/api/multisplit.cuh:199-313); keys and
    values of 32 or 64 bits (the reference's KeyT, ValueT)."""
    for name, t in (("d_key_in", d_key_in), ("d_value_in", d_value_in),
                    ("d_key_out", d_key_out), ("d_value_out", d_value_out)):
        _check_len(name, t, num_elements)
    h_off, _ = _offsets_buffer(multisplit_offset, context.num_buckets)
    if _is_wide(d_key_in, d_value_in, d_key_out, d_value_out):
        _multisplit_wide(d_key_in, d_value_in, d_key_out, d_value_out, num_elements, context,
                         bucket_identifier, h_off, None, _stream(d_key_in, d_key_out),
                         "multisplit_key_value")
        return
    _check(lib().gms_multisplit_key_value(context.handle, _ptr(d_key_in), _ptr(d_value_in),
                                          _ptr(d_key_out), _ptr(d_value_out), int(num_elements),
                                          bucket_identifier.desc(), h_off,
                                          _stream(d_key_in, d_key_out)), "multisplit_key_value")


def multisplit(d_key_in, d_key_out, num_elements: int, context: multisplit_context,
               bucket_identifier, d_value_in=None, d_value_out=None, d_offsets=None,
               stream=None) -> None:
    """Asynchronous form: optional values, bucket starts into device memory."""
    st = stream if stream is not None else _stream(d_key_in, d_key_out)
    if _is_wide(d_key_in, d_value_in, d_key_out, d_value_out):
        _multisplit_wide(d_key_in, d_value_in, d_key_out, d_value_out, num_elements, context,
                         bucket_identifier, None, _ptr(d_offsets), st, "gms_multisplit_wide")
        return
    _check(lib().gms_multisplit(context.handle, _ptr(d_key_in), _ptr(d_value_in), _ptr(d_key_out),
                                _ptr(d_value_out), int(num_elements), bucket_identifier.desc(),
                                _ptr(d_offsets), st), "gms_multisplit")


def multisplit_noalloc(d_key_in, d_key_out, num_elements: int, num_buckets: int,
                       bucket_identifier, d_value_in=None, d_value_out=None, d_offsets=None,
                       stream=None) -> None:
    """:func:`multisplit` without a context: allocates nothing; scratch is the
    library's per-GPU array of GMS_NOALLOC_SCRATCH_BYTES (32 MiB), shared by
    every call on that GPU, so calls that overlap in time on one GPU race.
    Up to 1024 buckets, 32-bit keys and values."""
    st = stream if stream is not None else _stream(d_key_in, d_key_out)
    _check(lib().gms_multisplit_noalloc(_ptr(d_key_in), _ptr(d_value_in), _ptr(d_key_out),
                                        _ptr(d_value_out), int(num_elements), int(num_buckets),
                                        bucket_identifier.desc(), _ptr(d_offsets), st),
           "gms_multisplit_noalloc")


def set_profiling(context: multisplit_context, enable: bool) -> None:
    _check(lib().gms_multisplit_set_profiling(context.handle, int(bool(enable))), "set_profiling")


def read_profile(context: multisplit_context):
    """(ms_prescan, ms_scan, ms_postscan, calls) summed since the last read."""
    a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    n = ctypes.c_uint32()
    _check(lib().gms_multisplit_profile(context.handle, ctypes.byref(a), ctypes.byref(b),
                                        ctypes.byref(c), ctypes.byref(n)), "profile")
    return a.value, b.value, c.value, n.value


def prescan_matrix(d_key_in, num_elements: int, context: multisplit_context, bucket_identifier,
                   d_matrix, key_value: bool = False) -> None:
    _check(lib().gms_multisplit_prescan_matrix(context.handle, _ptr(d_key_in), int(num_elements),
                                               bucket_identifier.desc(), int(bool(key_value)),
                                               _ptr(d_matrix), _stream(d_key_in)), "prescan_matrix")


def exclusive_scan(d_in, d_out, num_elements: int, context: multisplit_context) -> None:
    _check(lib().gms_exclusive_scan_u32(context.handle, _ptr(d_in), _ptr(d_out), int(num_elements),
                                        _stream(d_in)), "exclusive_scan")


# ---------------------------------------------------------------------------
# Multisplit radix sort (src/api/multisplit_sort.cuh, multisplit_sort_pairs.cuh)
# ---------------------------------------------------------------------------
class ms_sort_context:
    """ms_sort_context (src/api/multisplit_sort.cuh:100-116)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().gms_sort_create(ctypes.byref(h)), "ms_sort_context")
        self._h = h
        self.digit_bits = 0

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if self._h and _lib is not None:
                _lib.gms_sort_destroy(self._h)
                self._h = None
        except Exception:
            pass


def multisplit_sort_release_memory(context: ms_sort_context) -> None:
    """src/api/multisplit_sort.cuh:312-318."""
    _check(lib().gms_sort_release(context.handle), "multisplit_sort_release_memory")


def _sort_api(bits: int):
    def allocate(n_elements: int, context: ms_sort_context, key_bytes: int = 4) -> None:
        if key_bytes == 4:
            st = lib().gms_sort_allocate(context.handle, int(n_elements), bits, 0)
        else:
            st = lib().gms_sort_allocate_wide(context.handle, int(n_elements), bits, key_bytes, 0)
        _check(st, f"multisplit_sort_{bits}bit_allocate")
        context.digit_bits = bits

    def allocate_pairs(n_elements: int, context: ms_sort_context, key_bytes: int = 4,
                       value_bytes: int = 4) -> None:
        if key_bytes == 4 and value_bytes == 4:
            st = lib().gms_sort_allocate(context.handle, int(n_elements), bits, 1)
        else:
            st = lib().gms_sort_allocate_wide(context.handle, int(n_elements), bits, key_bytes,
                                              value_bytes)
        _check(st, f"multisplit_sort_{bits}bit_allocate_pairs")
        context.digit_bits = bits

    def sort(d_key_in, d_key_out, n_elements: int, context: ms_sort_context) -> None:
        _check_len("d_key_in", d_key_in, n_elements)
        _check_len("d_key_out", d_key_out, n_elements)
        if _is_wide(d_key_in, d_key_out):
            _check(lib().gms_sort_wide(context.handle, _ptr(d_key_in, (8,)), _ptr(d_key_out, (8,)),
                                       None, None, int(n_elements), 8, 0,
                                       _stream(d_key_in, d_key_out)), f"multisplit_sort_{bits}bit")
            return
        _check(lib().gms_sort(context.handle, _ptr(d_key_in), _ptr(d_key_out), None, None,
                              int(n_elements), bits, _stream(d_key_in, d_key_out)),
               f"multisplit_sort_{bits}bit")

    def sort_pairs(d_key_in, d_key_out, d_value_in, d_value_out, n_elements: int,
                   context: ms_sort_context) -> None:
        for name, t in (("d_key_in", d_key_in), ("d_key_out", d_key_out),
                        ("d_value_in", d_value_in), ("d_value_out", d_value_out)):
            _check_len(name, t, n_elements)
        if _is_wide(d_key_in, d_key_out, d_value_in, d_value_out):
            kb, vb = _elem_bytes(d_key_in, d_key_out), _elem_bytes(d_value_in, d_value_out)
            _check(lib().gms_sort_wide(context.handle, _ptr(d_key_in, (4, 8)), _ptr(d_key_out, (4, 8)),
                                       _ptr(d_value_in, (4, 8)), _ptr(d_value_out, (4, 8)),
                                       int(n_elements), kb, vb, _stream(d_key_in, d_key_out)),
                   f"multisplit_sort_{bits}bit_pairs")
            return
        _check(lib().gms_sort(context.handle, _ptr(d_key_in), _ptr(d_key_out), _ptr(d_value_in),
                              _ptr(d_value_out), int(n_elements), bits,
                              _stream(d_key_in, d_key_out)), f"multisplit_sort_{bits}bit_pairs")

    return allocate, allocate_pairs, sort, sort_pairs


(multisplit_sort_4bit_allocate, multisplit_sort_4bit_allocate_pairs,
 multisplit_sort_4bit, multisplit_sort_4bit_pairs) = _sort_api(4)
(multisplit_sort_5bit_allocate, multisplit_sort_5bit_allocate_pairs,
 multisplit_sort_5bit, multisplit_sort_5bit_pairs) = _sort_api(5)
(multisplit_sort_6bit_allocate, multisplit_sort_6bit_allocate_pairs,
 multisplit_sort_6bit, multisplit_sort_6bit_pairs) = _sort_api(6)
(multisplit_sort_7bit_allocate, multisplit_sort_7bit_allocate_pairs,
 multisplit_sort_7bit, multisplit_sort_7bit_pairs) = _sort_api(7)
(multisplit_sort_8bit_allocate, multisplit_sort_8bit_allocate_pairs,
 multisplit_sort_8bit, multisplit_sort_8bit_pairs) = _sort_api(8)


# ---------------------------------------------------------------------------
# Compaction (src/kernels/compaction/multisplit2_compaction.cuh:474-573)
# ---------------------------------------------------------------------------
class compaction_context:
    """compaction_context (multisplit2_compaction.cuh:477-489)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().gms_compaction_create(ctypes.byref(h)), "compaction_context")
        self._h = h

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if self._h and _lib is not None:
                _lib.gms_compaction_destroy(self._h)
                self._h = None
        except Exception:
            pass


def compaction_allocate_key_only(num_elements: int, context: compaction_context) -> None:
    _check(lib().gms_compaction_allocate(context.handle, int(num_elements), 0),
           "compaction_allocate_key_only")


def compaction_allocate_key_value(num_elements: int, context: compaction_context) -> None:
    _check(lib().gms_compaction_allocate(context.handle, int(num_elements), 1),
           "compaction_allocate_key_value")


def compaction_release_memory(context: compaction_context) -> None:
    _check(lib().gms_compaction_release(context.handle), "compaction_release_memory")


def _compaction(d_key_in, d_value_in, d_key_out, d_value_out, num_elements, context,
                bucket_identifier, h_count, d_count, what):
    st = _stream(d_key_in, d_key_out)
    if _is_wide(d_key_in, d_value_in, d_key_out, d_value_out):
        kb = _elem_bytes(d_key_in, d_key_out)
        vb = 0 if d_value_in is None and d_value_out is None else _elem_bytes(d_value_in, d_value_out)
        _check(lib().gms_compaction_wide(context.handle, _ptr(d_key_in, (4, 8)),
                                         _ptr(d_value_in, (4, 8)), _ptr(d_key_out, (4, 8)),
                                         _ptr(d_value_out, (4, 8)), int(num_elements), kb, vb,
                                         bucket_identifier.desc64(), h_count, d_count, st), what)
    else:
        _check(lib().gms_compaction(context.handle, _ptr(d_key_in), _ptr(d_value_in),
                                    _ptr(d_key_out), _ptr(d_value_out), int(num_elements),
                                    bucket_identifier.desc(), h_count, d_count, st), what)


def compaction_key_only(d_key_in, d_key_out, num_elements: int, context: compaction_context,
                        bucket_identifier) -> int:
    """Stable 2-bucket multisplit; returns the bucket-0 count (blocking).
    32- or 64-bit keys."""
    cnt = ctypes.c_uint32()
    _compaction(d_key_in, None, d_key_out, None, num_elements, context, bucket_identifier,
                ctypes.addressof(cnt), None, "compaction_key_only")
    return int(cnt.value)


def compaction_key_value(d_key_in, d_value_in, d_key_out, d_value_out, num_elements: int,
                         context: compaction_context, bucket_identifier) -> int:
    cnt = ctypes.c_uint32()
    _compaction(d_key_in, d_value_in, d_key_out, d_value_out, num_elements, context,
                bucket_identifier, ctypes.addressof(cnt), None, "compaction_key_value")
    return int(cnt.value)


def compaction_device_count(d_key_in, d_key_out, num_elements: int, context: compaction_context,
                            bucket_identifier, d_count, d_value_in=None, d_value_out=None) -> None:
    """Non-blocking compaction: the bucket-0 count goes to device memory."""
    _compaction(d_key_in, d_value_in, d_key_out, d_value_out, num_elements, context,
                bucket_identifier, None, _ptr(d_count), "compaction")


# ---------------------------------------------------------------------------
# Histogram (src/kernels/histogram/multisplit2_histograms.cuh)
# ---------------------------------------------------------------------------
def histogram(d_key_in, num_elements: int, d_histogram, num_bins: int, bucket_identifier) -> None:
    """Adds bucket-id counts of num_elements keys (32- or 64-bit) into
    d_histogram[num_bins]."""
    if _is_wide(d_key_in):
        _check(lib().gms_histogram_wide(_ptr(d_key_in, (8,)), int(num_elements), 8,
                                        _ptr(d_histogram), int(num_bins),
                                        bucket_identifier.desc64(), _stream(d_key_in)), "histogram")
        return
    _check(lib().gms_histogram_u32(_ptr(d_key_in), int(num_elements), _ptr(d_histogram),
                                   int(num_bins), bucket_identifier.desc(), _stream(d_key_in)),
           "histogram")


def multisplit2_histogram_even(d_key_in, num_elements: int, d_histogram, lower_level: float,
                               upper_level: float, inverse_delta: float, num_bins: int) -> None:
    """Even-bin float histogram, bin = (uint)((x - lower) * inverse_delta)."""
    _check(lib().gms_histogram_even_f32(_ptr(d_key_in), int(num_elements), _ptr(d_histogram),
                                        int(num_bins), float(lower_level), float(upper_level),
                                        float(inverse_delta), _stream(d_key_in)),
           "histogram_even")


def multisplit2_histogram_range(d_key_in, num_elements: int, d_histogram, d_levels,
                                num_levels: int) -> None:
    """Counts per custom bin [levels[i], levels[i+1]) added into d_histogram
    (num_levels - 1 bins); float32 samples/levels, or 32-bit integer ones
    read as uint32 (src/kernels/histogram/multisplit2_histograms.cuh:397-699)."""
    import torch
    is_float = hasattr(d_key_in, "dtype") and d_key_in.dtype == torch.float32
    fn = lib().gms_histogram_range_f32 if is_float else lib().gms_histogram_range_u32
    _check(fn(_ptr(d_key_in), int(num_elements), _ptr(d_histogram), int(num_levels),
              _ptr(d_levels), _stream(d_key_in)), "histogram_range")


def copy_u32(d_src, d_dst, n: int) -> None:
    _check(lib().gms_copy_u32(_ptr(d_src), _ptr(d_dst), int(n), _stream(d_src)), "copy_u32")


# ---------------------------------------------------------------------------
# Batched multisplit (no reference counterpart): many independent
# multisplits, planned once, run with a fixed number of launches, replayable
# through a hipGraph, tile shape picked by the autotuner.
# ---------------------------------------------------------------------------
class BatchItem:
    """One multisplit of a batch (device buffers as torch tensors or pointers)."""

    def __init__(self, d_key_in, d_key_out, num_elements: int, num_buckets: int,
                 bucket_identifier=None, d_value_in=None, d_value_out=None, d_offsets=None):
        self.d_key_in, self.d_key_out = d_key_in, d_key_out
        self.d_value_in, self.d_value_out = d_value_in, d_value_out
        self.n = int(num_elements)
        self.num_buckets = int(num_buckets)
        self.f = bucket_identifier if bucket_identifier is not None else range_bucket(num_buckets)
        self.d_offsets = d_offsets
        for name in ("d_key_in", "d_key_out", "d_value_in", "d_value_out"):
            _check_len(name, getattr(self, name), self.n)
        _check_len("d_offsets", d_offsets, self.num_buckets)

    def c_item(self) -> gms_batch_item:
        return gms_batch_item(_ptr(self.d_key_in), _ptr(self.d_value_in), _ptr(self.d_key_out),
                              _ptr(self.d_value_out), self.n, self.num_buckets, self.f.desc(),
                              _ptr(self.d_offsets))


class BatchSortItem:
    """One stable radix sort of a batch: the digit schedule of
    ``multisplit_sort_<digit_bits>bit[_pairs]`` (src/api/multisplit_sort.cuh:
    320-559), as :func:`sort`; 32- or 64-bit keys and values, from the
    tensors' element size.  The outputs double as a pass buffer; the input is
    never written."""

    def __init__(self, d_key_in, d_key_out, num_elements: int, digit_bits: int = 8,
                 d_value_in=None, d_value_out=None):
        self.d_key_in, self.d_key_out = d_key_in, d_key_out
        self.d_value_in, self.d_value_out = d_value_in, d_value_out
        self.n = int(num_elements)
        self.digit_bits = int(digit_bits)
        self.key_bytes = _elem_bytes(d_key_in, d_key_out)
        self.value_bytes = (_elem_bytes(d_value_in, d_value_out)
                            if d_value_in is not None or d_value_out is not None else 4)
        for name in ("d_key_in", "d_key_out", "d_value_in", "d_value_out"):
            _check_len(name, getattr(self, name), self.n)

    def c_item(self) -> gms_batch_sort_item:
        return gms_batch_sort_item(_ptr(self.d_key_in, (4, 8)), _ptr(self.d_value_in, (4, 8)),
                                   _ptr(self.d_key_out, (4, 8)), _ptr(self.d_value_out, (4, 8)),
                                   self.n, self.digit_bits, self.key_bytes, self.value_bytes)


class BatchWideItem:
    """One multisplit of a batch whose keys and/or values are 64-bit (int64
    tensors), as :func:`multisplit` on such tensors (the reference's
    KeyT/ValueT templates, src/api/multisplit.cuh:81-313); up to 256
    buckets, identifiers with their 64-bit parameter."""

    def __init__(self, d_key_in, d_key_out, num_elements: int, num_buckets: int,
                 bucket_identifier=None, d_value_in=None, d_value_out=None, d_offsets=None):
        self.d_key_in, self.d_key_out = d_key_in, d_key_out
        self.d_value_in, self.d_value_out = d_value_in, d_value_out
        self.n = int(num_elements)
        self.num_buckets = int(num_buckets)
        self.f = bucket_identifier if bucket_identifier is not None else range_bucket(num_buckets)
        self.d_offsets = d_offsets
        self.key_bytes = _elem_bytes(d_key_in, d_key_out)
        self.value_bytes = (_elem_bytes(d_value_in, d_value_out)
                            if d_value_in is not None or d_value_out is not None else 0)
        for name in ("d_key_in", "d_key_out", "d_value_in", "d_value_out"):
            _check_len(name, getattr(self, name), self.n)
        _check_len("d_offsets", d_offsets, self.num_buckets)

    def c_item(self) -> gms_batch_wide_item:
        return gms_batch_wide_item(_ptr(self.d_key_in, (4, 8)), _ptr(self.d_value_in, (4, 8)),
                                   _ptr(self.d_key_out, (4, 8)), _ptr(self.d_value_out, (4, 8)),
                                   self.n, self.num_buckets, self.key_bytes, self.value_bytes,
                                   self.f.desc64(), _ptr(self.d_offsets))


class Batch:
    """A planned batch of independent multisplits and radix sorts.

    ``run()`` launches it; ``capture()`` records the launches into a hipGraph
    that ``replay()`` relaunches; ``autotune()`` picks the tile shape and grid.
    The items' tensors are kept alive by this object.  Sorts run as passes of
    one batched multisplit per step: every multisplit and the first digit pass
    of every sort in step 0, the p-th pass of every sort in step p.
    """

    def __init__(self, items, stream=None, sorts=(), wide=()):
        h = ctypes.c_void_p()
        _check(lib().gms_batch_create(ctypes.byref(h)), "gms_batch_create")
        self._h = h
        self.items = list(items)
        self.sorts = list(sorts)
        self.wide = list(wide)
        arr = (gms_batch_item * max(len(self.items), 1))(*[it.c_item() for it in self.items])
        self._stream = stream
        if self.sorts or self.wide:
            sarr = (gms_batch_sort_item * max(len(self.sorts), 1))(
                *[s.c_item() for s in self.sorts])
            warr = (gms_batch_wide_item * max(len(self.wide), 1))(*[w.c_item() for w in self.wide])
            _check(lib().gms_batch_plan_wide(self._h, arr, len(self.items), warr, len(self.wide),
                                             sarr, len(self.sorts), self._s()),
                   "gms_batch_plan_wide")
        else:
            _check(lib().gms_batch_plan(self._h, arr, len(self.items), self._s()),
                   "gms_batch_plan")

    def _s(self):
        if self._stream is not None:
            return self._stream
        for it in self.items + self.sorts + self.wide:
            s = _stream(it.d_key_in, it.d_key_out)
            if s is not None:
                return s
        return None

    def run(self) -> None:
        _check(lib().gms_batch_run(self._h, self._s()), "gms_batch_run")

    def capture(self) -> None:
        _check(lib().gms_batch_capture(self._h), "gms_batch_capture")

    def replay(self) -> None:
        _check(lib().gms_batch_launch_graph(self._h, self._s()), "gms_batch_launch_graph")

    def autotune(self, reps: int = 3):
        t = gms_batch_tuning()
        ms = ctypes.c_float()
        _check(lib().gms_batch_autotune(self._h, int(reps), self._s(), ctypes.byref(t),
                                        ctypes.byref(ms)), "gms_batch_autotune")
        return {"items_per_thread": t.items_per_thread, "persistent": t.persistent,
                "ms_per_run": float(ms.value)}

    @property
    def tuning(self):
        t = gms_batch_tuning()
        _check(lib().gms_batch_get_tuning(self._h, ctypes.byref(t)), "gms_batch_get_tuning")
        return {"items_per_thread": t.items_per_thread, "persistent": t.persistent}

    @tuning.setter
    def tuning(self, value):
        t = gms_batch_tuning(int(value["items_per_thread"]), int(value["persistent"]))
        _check(lib().gms_batch_set_tuning(self._h, t), "gms_batch_set_tuning")

    def stats(self):
        k, m = ctypes.c_uint64(), ctypes.c_uint64()
        w, l = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().gms_batch_stats(self._h, ctypes.byref(k), ctypes.byref(w), ctypes.byref(m),
                                     ctypes.byref(l)), "gms_batch_stats")
        return {"keys": k.value, "work_items": w.value, "matrix_words": m.value,
                "launches": l.value}

    def __del__(self):
        try:
            if self._h and _lib is not None:
                _lib.gms_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Request server with continuous batching (no reference counterpart).
# ---------------------------------------------------------------------------
OP_MULTISPLIT, OP_SORT = 0, 1


def _request_widths(d_key_in, d_key_out, d_value_in, d_value_out):
    kb = _elem_bytes(d_key_in, d_key_out)
    vb = (_elem_bytes(d_value_in, d_value_out)
          if d_value_in is not None or d_value_out is not None else 4)
    return kb, vb


def multisplit_request(d_key_in, d_key_out, num_elements: int, num_buckets: int,
                       bucket_identifier=None, d_value_in=None, d_value_out=None,
                       d_offsets=None) -> gms_request:
    """A multisplit request; int64 keys or values make it a 64-bit one (the
    identifier's parameter may then need 64 bits)."""
    f = bucket_identifier if bucket_identifier is not None else range_bucket(num_buckets)
    for name, t in (("d_key_in", d_key_in), ("d_key_out", d_key_out),
                    ("d_value_in", d_value_in), ("d_value_out", d_value_out)):
        _check_len(name, t, num_elements)
    _check_len("d_offsets", d_offsets, num_buckets)
    kb, vb = _request_widths(d_key_in, d_key_out, d_value_in, d_value_out)
    wide = kb == 8 or (vb == 8 and d_value_in is not None)
    desc = gms_bucket_desc(0, 0, 0) if wide else f.desc()
    return gms_request(OP_MULTISPLIT, _ptr(d_key_in, (4, 8)), _ptr(d_value_in, (4, 8)),
                       _ptr(d_key_out, (4, 8)), _ptr(d_value_out, (4, 8)), int(num_elements),
                       int(num_buckets), desc, 0, _ptr(d_offsets), kb, vb, f.desc64())


def sort_request(d_key_in, d_key_out, num_elements: int, digit_bits: int = 8, d_value_in=None,
                 d_value_out=None) -> gms_request:
    """A radix sort request; int64 keys or values make it a 64-bit one."""
    for name, t in (("d_key_in", d_key_in), ("d_key_out", d_key_out),
                    ("d_value_in", d_value_in), ("d_value_out", d_value_out)):
        _check_len(name, t, num_elements)
    kb, vb = _request_widths(d_key_in, d_key_out, d_value_in, d_value_out)
    return gms_request(OP_SORT, _ptr(d_key_in, (4, 8)), _ptr(d_value_in, (4, 8)),
                       _ptr(d_key_out, (4, 8)), _ptr(d_value_out, (4, 8)), int(num_elements), 0,
                       gms_bucket_desc(0, 0, 0), int(digit_bits), None, kb, vb,
                       gms_bucket_desc64(0, 0, 0))


class Server:
    """Continuous-batching request server on the current device.

    ``submit`` queues a request built by ``multisplit_request`` or
    ``sort_request`` and returns a ticket; ``wait`` blocks until it is done
    and returns its latency in microseconds (submit to completion).  The
    caller keeps the request's buffers alive until then.
    """

    def __init__(self, scratch_budget_bytes: int = 64 << 30, max_inflight_steps: int = 2,
                 max_step_keys: int = 0, items_per_thread: int = 24):
        cfg = gms_server_config(int(scratch_budget_bytes), int(max_inflight_steps),
                                int(max_step_keys), int(items_per_thread))
        h = ctypes.c_void_p()
        _check(lib().gms_server_create(ctypes.byref(cfg), ctypes.byref(h)), "gms_server_create")
        self._h = h

    def submit(self, req: gms_request) -> int:
        t = ctypes.c_uint64()
        _check(lib().gms_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)),
               "gms_server_submit")
        return int(t.value)

    def poll(self, ticket: int):
        d = ctypes.c_int()
        lat = ctypes.c_double()
        _check(lib().gms_server_poll(self._h, int(ticket), ctypes.byref(d), ctypes.byref(lat)),
               "gms_server_poll")
        return bool(d.value), float(lat.value)

    def wait(self, ticket: int) -> float:
        lat = ctypes.c_double()
        _check(lib().gms_server_wait(self._h, int(ticket), ctypes.byref(lat)), "gms_server_wait")
        return float(lat.value)

    def stats(self):
        st = gms_server_stats()
        _check(lib().gms_server_get_stats(self._h, ctypes.byref(st)), "gms_server_get_stats")
        return {k: int(getattr(st, k)) for k, _ in gms_server_stats._fields_}

    def close(self) -> None:
        if self._h:
            _check(lib().gms_server_destroy(self._h), "gms_server_destroy")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if self._h and _lib is not None:
                _lib.gms_server_destroy(self._h)
                self._h = None
        except Exception:
            pass
