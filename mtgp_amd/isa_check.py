"""Guard against a code-generation hazard seen with this toolchain.

The gfx950 backend once emitted `s_cselect_b*` (which selects on SCC)
directly after a 64-bit scalar subtract whose borrow had overwritten SCC,
dropping the `s_and_b64 ..., vcc, exec` that should have moved a VALU
compare result into SCC.  The select then always took one side: a tail
step computed with the wrong length.  This check scans the generated
assembly for any `s_cselect` whose SCC was last written by an add/subtract
with carry in the same basic block and fails the build if it finds one.
"""
import re
import sys

_SCC_WRITER = re.compile(
    r"^\s+(s_cmp|s_cmpk|s_and_|s_or_|s_xor_|s_andn2|s_orn2|s_nand|s_nor|s_xnor|s_bitcmp|"
    r"s_add_u32|s_addc_u32|s_sub_u32|s_subb_u32|s_add_i32|s_sub_i32|s_lshl|s_lshr|s_ashr|"
    r"s_bfe|s_not|s_min|s_max|s_abs|s_bcnt|s_ff|s_and_saveexec|s_or_saveexec|s_andn2_saveexec)")
_CARRY = re.compile(r"^\s+s_(subb_u32|addc_u32|sub_u32|add_u32)")


def scan(asm_text):
    """Returns a list of (line number, writer, select) suspects."""
    out, last = [], None
    for i, ln in enumerate(asm_text.splitlines()):
        if ln.startswith(".LBB") or (ln and not ln[0].isspace() and ln.endswith(":")):
            last = None
        if "s_cselect" in ln and last is not None and _CARRY.match(last):
            out.append((i + 1, last.strip(), ln.strip()))
        if _SCC_WRITER.match(ln):
            last = ln
    return out


if __name__ == "__main__":
    bad = scan(open(sys.argv[1]).read())
    for b in bad:
        print("line %d: %s -> %s" % b)
    sys.exit(1 if bad else 0)
