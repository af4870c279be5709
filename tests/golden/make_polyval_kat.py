"""Extract the POLYVAL repeated-squaring known-answer table as data.

The reference holds a constant table (p3/src/polyval/portable.rs:169-376)
with entries ``T[i][k][j] = from_canonical(j << 4k) ^ (2 ^ (2 ^ (i + 1)))``
in Montgomery form, i in 0..6, k in 0..32, j in 0..16.  Those 3,072 values
pin ``from_canonical`` (the R^2 constant) and the Montgomery multiply
(repeated squaring) bit-exactly.  This script reads the values out of the
source text and writes them, one hex value per line in [i][k][j] order, to
``polyval_nibble_pow_table.txt`` next to this file.  Only the numbers are
kept; no source text is copied.

Run from the repository root:  python tests/golden/make_polyval_kat.py
"""
import pathlib
import re

SRC = pathlib.Path("/root/reference/p3/src/polyval/portable.rs")
OUT = pathlib.Path(__file__).with_name("polyval_nibble_pow_table.txt")


def main() -> None:
    text = SRC.read_text()
    start = text.index("POLYVAL_NIBBLE_POW_2_N_TABLE: [[[u128; 16]; 32]; 6] = [")
    body = text[start:]
    body = body[body.index("= [") + 3:]
    values = re.findall(r"0x[0-9a-fA-F]+", body)
    assert len(values) == 6 * 32 * 16, len(values)
    lines = ["# T[i][k][j] = from_canonical(j << 4k)^(2^(2^(i+1))), Montgomery form, "
             "order i(6) k(32) j(16); source p3/src/polyval/portable.rs:169-376"]
    lines += [f"{int(v, 16):032x}" for v in values]
    OUT.write_text("\n".join(lines) + "\n")
    print(f"wrote {len(values)} values to {OUT}")


if __name__ == "__main__":
    main()
