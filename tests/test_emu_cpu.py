"""CPU parity of the library's HIP kernel sources, built for the host emulation
(tests/emu): the same prescan/scan/postscan/sort/histogram code runs with one
host thread per GPU thread and is compared with the oracle at small sizes.

This pins the kernels' index arithmetic and stability on the CPU; the GPU
tests pin the compiled gfx950 code.  The AddressSanitizer build of the same
sources is `make -C tests/emu emu_check_asan && tests/emu/emu_check_asan`.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import gpumultisplit_amd as gms
import oracle
from gpumultisplit_amd import workloads as W

EMU_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")
EMU_LIB = os.path.join(EMU_DIR, "libgpumultisplit_emu.so")


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", EMU_DIR, "libgpumultisplit_emu.so"], check=True)
    L = ctypes.CDLL(EMU_LIB)
    for name, res, args in gms._SIGS:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def p(a):
    return a.ctypes.data if a is not None else None


def emu_multisplit(L, keys, B, functor, kv):
    n = keys.size
    h = ctypes.c_void_p()
    assert L.gms_multisplit_create(B, ctypes.byref(h)) == 0
    try:
        assert L.gms_multisplit_allocate(h, n, int(kv)) == 0
        vals = np.arange(n, dtype=np.uint32)
        kout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        vout = np.full(max(n, 1), 0x5A5A5A5A, np.uint32)
        offs = np.zeros(B, np.uint32)
        if kv:
            st = L.gms_multisplit_key_value(h, p(keys), p(vals), p(kout), p(vout), n,
                                            functor.desc(), p(offs), None)
        else:
            st = L.gms_multisplit_key_only(h, p(keys), p(kout), n, functor.desc(), p(offs), None)
        assert st == 0
        return kout[:n], (vout[:n] if kv else None), offs
    finally:
        L.gms_multisplit_destroy(h)


@pytest.mark.parametrize("B", [1, 2, 3, 8, 17, 64, 128, 256])
@pytest.mark.parametrize("kv", [False, True])
def test_emulated_multisplit_matches_oracle(emu, B, kv):
    tile = gms.subproblem_size_bms_key_value(256)
    for n in (0, 1, 65, tile, tile + 1):
        keys = W.uniform_keys(n, seed=B * 7 + n)
        got = emu_multisplit(emu, keys, B, gms.range_bucket(B), kv)
        wk, wv, wo = oracle.multisplit(keys, B, oracle.RANGE, 0,
                                       np.arange(n, dtype=np.uint32) if kv else None)
        np.testing.assert_array_equal(got[0], wk)
        np.testing.assert_array_equal(got[2], wo)
        if kv:
            np.testing.assert_array_equal(got[1], wv)


def test_emulated_skewed_and_delta(emu):
    tile = gms.subproblem_size_bms_key_value(256)
    n = 2 * tile + 3
    keys = W.keys_for("onehot90", n, 256, seed=5)
    got = emu_multisplit(emu, keys, 256, gms.range_bucket(256), True)
    wk, wv, wo = oracle.multisplit(keys, 256, oracle.RANGE, 0, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(got[0], wk)
    np.testing.assert_array_equal(got[1], wv)
    delta = (n + 31) // 32
    keys = W.delta_keys(n, 32, delta, seed=6)
    got = emu_multisplit(emu, keys, 32, gms.delta_func(delta), True)
    wk, wv, _ = oracle.multisplit(keys, 32, oracle.DELTA, delta, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(got[0], wk)
    np.testing.assert_array_equal(got[1], wv)


@pytest.mark.parametrize("bits", [7, 8])
def test_emulated_sort(emu, bits):
    n = 5000
    keys = W.uniform_keys(n, seed=bits)
    keys[::5] = keys[1]
    vals = np.arange(n, dtype=np.uint32)
    kout, vout = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    h = ctypes.c_void_p()
    assert emu.gms_sort_create(ctypes.byref(h)) == 0
    assert emu.gms_sort_allocate(h, n, bits, 1) == 0
    assert emu.gms_sort(h, p(keys), p(kout), p(vals), p(vout), n, bits, None) == 0
    emu.gms_sort_destroy(h)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(kout, keys[order])
    np.testing.assert_array_equal(vout, vals[order])
