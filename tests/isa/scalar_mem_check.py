"""Scans gfx950 assembly for writes through the scalar data cache (scalar
memory stores, scalar atomics, and the scalar cache write-back/discard
instructions).  The library's device code writes memory with vector
instructions only; this check keeps it that way.  CPU-side checker, never run
on a GPU.  Usage: python scalar_mem_check.py file.s [...]; exit 1 on a hit."""
import re
import sys

PATTERN = re.compile(r"^\s*(s_store_\w+|s_buffer_store_\w+|s_scratch_store_\w+|s_atomic_\w+|"
                     r"s_buffer_atomic_\w+|s_dcache_wb\w*|s_dcache_discard\w*)\b")
if __name__ != "__main__":
    raise ImportError("command-line tool")
hits = 0
for path in sys.argv[1:]:
    for ln, line in enumerate(open(path), 1):
        if PATTERN.match(line):
            hits += 1
            if hits <= 20:
                print(f"{path}:{ln}: {line.strip()}")
print("scalar-cache writes:", hits)
sys.exit(1 if hits else 0)
