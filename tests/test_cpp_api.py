"""The reference-compatible C++ API (include/gpumultisplit.hpp) driven by a C++
program written like the reference's drivers (tests/cpp/api_driver.cpp):
multisplit with delta_func and h_offsets, 7/8-bit sorts (keys and pairs,
checked against std::stable_sort), compaction with test_bucket<1>, and the
error behaviour.  The CPU test runs it on the emulation build of the kernels;
the GPU test runs it on libgpumultisplit.so.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def _build():
    subprocess.run(["make", "-s", "-C", CPP], check=True)


def test_header_compiles_standalone():
    # the header needs nothing but the C header (no HIP headers)
    src = '#include "gpumultisplit.hpp"\nint main() { return subproblem_size_bms_key_only(8) ? 0 : 0; }\n'
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I", os.path.join(os.path.dirname(HERE), "include"), "-x", "c++", "-"],
                       input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr


def test_cpp_api_driver_on_emulation():
    _build()
    r = subprocess.run([os.path.join(CPP, "bin", "api_driver_host"), "7000"], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0 and "API DRIVER OK" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 6145, 1 << 22])
def test_cpp_api_driver_on_gpu(dev, n):
    exe = os.path.join(CPP, "bin", "api_driver")
    assert os.path.exists(exe), "tests/cpp/bin/api_driver is not built"
    print(f"[case] api_driver n={n}", flush=True)
    r = subprocess.run([exe, str(n)], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0 and "API DRIVER OK" in r.stdout, r.stdout + r.stderr


def test_functor_driver_on_emulation():
    """Programmer-supplied functors through gpumultisplit_device.hpp on the
    CPU emulation of the kernels: multiplicative hash (2..1024 buckets), ids
    past the bucket count (clamped), popcount mod B (any B) and its parity as
    a compaction predicate, 64-bit keys/values with a 64-bit hash; every
    result against std::stable_sort of the positions by host-computed ids."""
    subprocess.run(["make", "-s", "-C", CPP, "bin/functor_driver_host"], check=True)
    r = subprocess.run([os.path.join(CPP, "bin", "functor_driver_host"), "0", "1", "6143", "6145",
                        "20000"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and "FUNCTOR DRIVER OK" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("n", ["0 1 6143 6144 6145 12289", "100003", str((1 << 22) + 7)])
def test_functor_driver_on_gpu(dev, n):
    """The same functor cases on the GPU: the driver's own instantiations of
    the library's kernel templates (compiled by hipcc in tests/cpp), linked
    to libgpumultisplit.so for the stages it runs."""
    exe = os.path.join(CPP, "bin", "functor_driver")
    assert os.path.exists(exe), "tests/cpp/bin/functor_driver is not built"
    print(f"[case] functor_driver n={n}", flush=True)
    r = subprocess.run([exe] + n.split(), capture_output=True, text=True, timeout=110)
    assert r.returncode == 0 and "FUNCTOR DRIVER OK" in r.stdout, r.stdout + r.stderr
