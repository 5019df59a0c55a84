"""Writes the small-field Nastran triangle meshes (8-character fields) of
tests/golden/decks:

  - tool_mini.nas: exponents written in the short forms the format allows
    ("5.0-4", "1.0+0", "-2.5-1", "1.00D0"); the expected values are in the
    comment line of the file itself;
  - press_plate.nas: a 5 x 5 mm plate at z = 0 of 3 x 3 nodes (8 triangles,
    n0n1 x n0n2 = -z, facing a workpiece below it), node coordinates in the
    implicit-exponent form ("-1.13-3" = -1.13e-3 m).  Its node lines are
    unevenly spaced so that no lattice point of a dx = 0.6 mm box lies on a
    triangle edge."""
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


# press plate: node lines in x and y (mm), written as implicit exponents
PX = ["-1.13-3", "1.41-3", "3.95-3"]
PY = ["-1.07-3", "1.37-3", "4.02-3"]


def press_plate():
    rows, tris = [], []
    nid = lambda i, j: 100 + 3 * j + i
    for j in range(3):
        for i in range(3):
            rows.append((nid(i, j), PX[i], PY[j], "0.0"))
    e = 1
    for j in range(2):
        for i in range(2):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            tris += [(e, a, c, b), (e + 1, a, d, c)]  # both -z
            e += 2
    return rows, tris


def write(path, title, rows, tris):
    with open(path, "w") as f:
        f.write(f"$ {title}\n")
        f.write("BEGIN BULK\n")
        for i, x, y, z in rows:
            f.write("GRID    " + f8(str(i)) + f8("") + f8(x) + f8(y) + f8(z) + "\n")
        for e, a, b, c in tris:
            f.write("CTRIA3  " + f8(str(e)) + f8("1") + f8(str(a)) + f8(str(b)) + f8(str(c))
                    + "\n")
        f.write("ENDDATA\n")


def main():
    decks = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks")
    write(os.path.join(decks, "tool_mini.nas"),
          "unit square (2 triangles) at z = 5e-4 and a tilted triangle", ROWS, TRIS)
    rows, tris = press_plate()
    write(os.path.join(decks, "press_plate.nas"),
          "5 x 5 mm plate at z = 0, 8 triangles with normal -z", rows, tris)


if __name__ == "__main__":
    main()
