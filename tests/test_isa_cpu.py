"""Compiles every HIP source of the library to gfx950 assembly (on the CPU) and
checks it for the condition-flag miscompile pattern (tests/isa/scc_check.py)
and for writes through the scalar data cache (tests/isa/scalar_mem_check.py)."""
import glob
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gpumultisplit_amd", "csrc")
CHECK = os.path.join(ROOT, "tests", "isa", "scc_check.py")
SMEM_CHECK = os.path.join(ROOT, "tests", "isa", "scalar_mem_check.py")


def test_device_code_isa_checks():
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with tempfile.TemporaryDirectory() as tmp:
        def compile_one(src):
            out = os.path.join(tmp, os.path.basename(src) + ".s")
            subprocess.run(["/opt/rocm/bin/hipcc", "-O3", "-std=c++17", "--offload-arch=gfx950",
                            "--cuda-device-only", "-S", "-I", os.path.join(ROOT, "include"),
                            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                            src, "-o", out], check=True, capture_output=True)
            return out
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
            outs = list(ex.map(compile_one, srcs))
        r = subprocess.run([sys.executable, CHECK] + outs, capture_output=True, text=True)
        assert r.returncode == 0, r.stdout[-4000:]
        r = subprocess.run([sys.executable, SMEM_CHECK] + outs, capture_output=True, text=True)
        assert r.returncode == 0, r.stdout[-4000:]
