"""Scans gfx950 assembly for SCC consumers (s_cselect, s_cbranch_scc*,
s_cmov) whose nearest preceding SCC producer, in program order through
fall-through code, is an arithmetic instruction (its SCC is a carry or an
overflow, not a condition).

That is the shape of a miscompile seen with a uniform 64-bit unsigned compare
(no scalar instruction for it on gfx950): the compare went to the VALU while
the select read SCC, left behind by an address add, so a ragged tile counted
as full.  The library avoids the construct; this check keeps it out.
Usage: python scc_check.py file.s [...]; prints findings, exit 1 if any."""
import re, sys
CONSUMERS = re.compile(r"^\s*(s_cselect_b(32|64)|s_cbranch_scc[01]|s_cmov_b(32|64))\b")
ARITH = re.compile(r"^\s*(s_add_u32|s_addc_u32|s_sub_u32|s_subb_u32|s_add_i32|s_sub_i32|s_addk_i32|s_absdiff_i32)\b")
PRODUCERS = re.compile(r"^\s*(s_cmp\w*|s_bitcmp\w*|s_and_b\d+|s_or_b\d+|s_xor_b\d+|s_andn2\w*|s_orn2\w*|s_nand\w*|s_nor\w*|s_xnor\w*|s_not_b\d+|s_min_\w+|s_max_\w+|s_lshl\w*|s_lshr\w*|s_ashr\w*|s_bfe\w*|s_abs_i32|s_and_saveexec\w*|s_or_saveexec\w*|s_andn2_saveexec\w*|s_cmpk\w*|s_add_u32|s_addc_u32|s_sub_u32|s_subb_u32|s_add_i32|s_sub_i32|s_addk_i32|s_absdiff_i32|s_bcnt\w*|s_quadmask\w*|s_wqm\w*|s_ff1\w*)\b")
def check(path):
    bad = []
    func = None; last = None
    for ln, line in enumerate(open(path), 1):
        if re.match(r"^_Z\w+:", line): func = line.split(":")[0]; last = None; continue
        if re.match(r"^\s*s_(branch|cbranch)", line) and not CONSUMERS.match(line):
            pass  # control flow does not write SCC
        if CONSUMERS.match(line):
            if last is not None and ARITH.match(last[1]) and not line.strip().startswith("s_cbranch"):
                bad.append((func, ln, last[1].strip(), line.strip()))
            elif last is not None and ARITH.match(last[1]) and line.strip().startswith("s_cbranch_scc"):
                bad.append((func, ln, last[1].strip(), line.strip()))
        if PRODUCERS.match(line):
            last = (ln, line)
    return bad
if __name__ != "__main__":
    raise ImportError("command-line tool")
tot = 0
for p in sys.argv[1:]:
    b = check(p)
    tot += len(b)
    for f, ln, prod, cons in b[:20]:
        print(f"{p}:{ln} {f[:70]}\n    SCC from: {prod}\n    used by:  {cons}")
print("suspicious:", tot)
sys.exit(1 if tot else 0)
