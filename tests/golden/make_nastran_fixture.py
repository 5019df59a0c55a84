"""Writes tests/golden/decks/tool_mini.nas: a small-field Nastran triangle
mesh (8-character fields) with exponents written in the short forms the
format allows ("5.0-4", "1.0+0", "-2.5-1", "1.00D0").  The expected values
are in the comment lines of the file itself."""
import os

ROWS = [  # (id, x, y, z) as 8-character field texts
    (10, "0.0", "0.0", "5.0-4"),
    (11, "1.0", "0.0", "5.0-4"),
    (12, "1.0", "1.0", "5.0-4"),
    (13, "0.0", "1.0", "0.00050"),
    (20, "-2.5-1", "1.0+0", "-1.0+0"),
    (21, "2.5E-01", "1.0", "-1.00D0"),
    (22, "0.0", "2.0", "0.0"),
]
TRIS = [(1, 10, 11, 12), (2, 10, 12, 13), (3, 20, 21, 22)]


def f8(s):
    assert len(s) <= 8
    return s.rjust(8)


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks", "tool_mini.nas")
    with open(out, "w") as f:
        f.write("$ unit square (2 triangles) at z = 5e-4 and a tilted triangle\n")
        f.write("BEGIN BULK\n")
        for i, x, y, z in ROWS:
            f.write("GRID    " + f8(str(i)) + f8("") + f8(x) + f8(y) + f8(z) + "\n")
        for e, a, b, c in TRIS:
            f.write("CTRIA3  " + f8(str(e)) + f8("1") + f8(str(a)) + f8(str(b)) + f8(str(c))
                    + "\n")
        f.write("ENDDATA\n")


if __name__ == "__main__":
    main()
