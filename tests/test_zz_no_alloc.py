"""GPU parity of gms_multisplit_noalloc (the multisplit that allocates nothing,
its scratch a per-GPU device array in the library) against gms_multisplit
with a context, both called through the C ABI: bit-exact keys, values and
bucket starts on two calls in a row, the second smaller than the first, so
that the second runs over the first call's leftover scratch.
"""
import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W
from gpu_util import banner, filled_u32, from_dev, to_dev

pytestmark = pytest.mark.gpu

# (num_buckets, key_value, first n, second n, key distribution)
CASES = [
    (256, True, 1 << 27, 1_000_003, "uniform"),
    (1024, False, 3_000_001, 6145, "zipf"),
    (5, True, 200_000, 1, "uniform"),
    (64, True, 1 << 20, 0, "onehot90"),
]


def _keys(dist, n, B, seed, dev):
    import torch
    if dist == "uniform" and n >= (1 << 24):
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        return torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32, device=dev,
                             generator=g)
    keys = W.uniform_keys(n, seed=seed) if dist == "uniform" else W.keys_for(dist, n, B, seed=seed)
    return to_dev(keys, dev) if n else filled_u32(0, dev)


@pytest.mark.parametrize("B,kv,n1,n2,dist", CASES)
def test_noalloc_equals_context_call(dev, B, kv, n1, n2, dist):
    import torch
    L = gms.lib()
    desc = gms.range_bucket(B).desc()
    ctx = gms.multisplit_context(B)
    (gms.multisplit_allocate_key_value if kv else gms.multisplit_allocate_key_only)(n1, ctx)
    stream = torch.cuda.current_stream().cuda_stream
    try:
        for call, n in enumerate((n1, n2)):
            banner(f"noalloc B={B} kv={kv} n={n} {dist} (call {call + 1} of 2)")
            kin = _keys(dist, n, B, 1000 * call + B, dev)
            vin = torch.arange(max(n, 1), dtype=torch.int32, device=dev) if kv else None
            outs = []
            for noalloc in (False, True):
                kout = filled_u32(n, dev)
                vout = filled_u32(n, dev) if kv else None
                offs = filled_u32(B, dev, 0xFFFFFFFF)
                args = (kin.data_ptr(), vin.data_ptr() if kv else None, kout.data_ptr(),
                        vout.data_ptr() if kv else None, n)
                if noalloc:
                    st = L.gms_multisplit_noalloc(*args, B, desc, offs.data_ptr(), stream)
                else:
                    st = L.gms_multisplit(ctx.handle, *args, desc, offs.data_ptr(), stream)
                assert st == 0, L.gms_status_string(st).decode()
                outs.append((kout, vout, offs))
            torch.cuda.synchronize()
            (k0, v0, o0), (k1, v1, o1) = outs
            assert torch.equal(k1[:n], k0[:n]), f"keys differ (n={n})"
            assert torch.equal(o1, o0), f"bucket starts differ (n={n})"
            if kv:
                assert torch.equal(v1[:n], v0[:n]), f"values differ (n={n})"
            # and both are the multisplit: the oracle at the sizes it finishes
            # in seconds, ascending bucket starts from 0 at every size
            o = from_dev(o1).astype(np.int64)
            assert n == 0 or (o[0] == 0 and np.all(np.diff(o) >= 0) and o[-1] <= n)
            if 0 < n <= (1 << 22):
                keys = from_dev(kin)[:n]
                wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0,
                                               np.arange(n, dtype=np.uint32) if kv else None)
                np.testing.assert_array_equal(from_dev(k1)[:n], wk)
                np.testing.assert_array_equal(from_dev(o1), wo)
                if kv:
                    np.testing.assert_array_equal(from_dev(v1)[:n], wv)
    finally:
        gms.multisplit_release_memory(ctx)


def test_noalloc_rejects_what_does_not_fit(dev):
    """Rejected on the host before any launch: more buckets than one pass
    takes, a matrix beyond the 32 MiB array, aliased outputs."""
    L = gms.lib()
    a = filled_u32(64, dev)
    b = filled_u32(64, dev)
    ptr_a, ptr_b = a.data_ptr(), b.data_ptr()
    assert L.gms_multisplit_noalloc(ptr_a, None, ptr_b, None, 64, 1025,
                                    gms.range_bucket(1025).desc(), None, None) == 2
    assert L.gms_multisplit_noalloc(ptr_a, None, ptr_b, None, (1 << 31), 1024,
                                    gms.range_bucket(1024).desc(), None, None) == 3
    assert L.gms_multisplit_noalloc(ptr_a, None, ptr_a + 64, None, 64, 4,
                                    gms.range_bucket(4).desc(), None, None) == 7
