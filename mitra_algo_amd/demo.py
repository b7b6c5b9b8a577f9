"""Small demonstration meshes for the detector (the generator behind the
reference's examples/*.off, tools/make_demo.py:8-56), and an OFF writer.

shape(level): the octahedron refined `level` times by splitting each triangle
at its edge midpoints pushed out to the unit sphere (new vertices numbered in
the order the faces first reach each edge, faces (a, ab, ca), (ab, b, bc),
(ca, bc, c), (ab, bc, ca)), then warped by
    (x, y, z) -> (0.7x + 0.11yz + 0.09z^2, 0.48y + 0.07x^2 + 0.05zx,
                  z + 0.12xy + 0.08x)
so that principal curvatures differ and no mirror is built in.
demo(kind): that shape moved to x - 1.6, plus
  reflection  its mirror image in x = 0 (faces rewound),
  rigid       a copy turned 0.65 rad about z and moved by (3.2, 0.3, 0.2),
  partial     the mirror pair and a decoy copy scaled 1.3 at (0.4, 4, 0.5),
  single      nothing else.
The arithmetic is written term by term in the order above so the files come
out byte-identical to the reference's examples.
"""
from __future__ import annotations

import argparse
import math
from pathlib import Path

_OCTAHEDRON_V = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
_OCTAHEDRON_F = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5),
                 (0, 3, 5)]


def _neumaier_sum(values) -> float:
    """Compensated sum (A. Neumaier, ZAMM 54, 1974), the float summation of
    Python's built-in sum() from 3.12 on, under which the reference's example
    files were written; spelled out so every interpreter gives them."""
    total = 0.0
    comp = 0.0
    for v in values:
        t = total + v
        if abs(total) >= abs(v):
            comp += (total - t) + v
        else:
            comp += (v - t) + total
        total = t
    return total + comp


def _refine(verts: list, faces: list) -> list:
    index = {}

    def mid(a: int, b: int) -> int:
        key = (a, b) if a < b else (b, a)
        if key not in index:
            s = [verts[a][k] + verts[b][k] for k in range(3)]
            r = math.sqrt(_neumaier_sum(c * c for c in s))
            index[key] = len(verts)
            verts.append(tuple(c / r for c in s))
        return index[key]

    out = []
    for a, b, c in faces:
        ab = mid(a, b)
        bc = mid(b, c)
        ca = mid(c, a)
        out.extend([(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)])
    return out


def shape(level: int = 2):
    verts = list(_OCTAHEDRON_V)
    faces = list(_OCTAHEDRON_F)
    for _ in range(level):
        faces = _refine(verts, faces)
    warped = [(0.7 * x + 0.11 * y * z + 0.09 * z * z,
               0.48 * y + 0.07 * x * x + 0.05 * z * x,
               z + 0.12 * x * y + 0.08 * x) for x, y, z in verts]
    return warped, faces


def demo(kind: str = "reflection", level: int = 2):
    verts, faces = shape(level)
    base = [(x - 1.6, y, z) for x, y, z in verts]
    n = len(base)
    if kind == "single":
        return base, faces
    if kind in ("reflection", "partial"):
        other = [(-x, y, z) for x, y, z in base]
        other_faces = [(a, c, b) for a, b, c in faces]
    elif kind == "rigid":
        c, s = math.cos(0.65), math.sin(0.65)
        other = [(c * x - s * y + 3.2, s * x + c * y + 0.3, z + 0.2) for x, y, z in base]
        other_faces = faces
    else:
        raise ValueError(kind)
    pair_v = base + other
    pair_f = faces + [(a + n, b + n, c + n) for a, b, c in other_faces]
    if kind != "partial":
        return pair_v, pair_f
    m = len(pair_v)
    decoy = [(1.3 * x + .4, 1.3 * y + 4, 1.3 * z + .5) for x, y, z in verts]
    return pair_v + decoy, pair_f + [(a + m, b + m, c + m) for a, b, c in faces]


def write_off(path, vertices, faces) -> None:
    """OFF with 17 significant digits per coordinate."""
    lines = [f"OFF\n{len(vertices)} {len(faces)} 0\n"]
    lines += [" ".join(format(c, ".17g") for c in p) + "\n" for p in vertices]
    lines += ["3 " + " ".join(str(i) for i in f) + "\n" for f in faces]
    Path(path).write_text("".join(lines))


def main() -> None:
    ap = argparse.ArgumentParser(description="write the demonstration meshes as OFF files")
    ap.add_argument("--output", type=Path, default=Path("examples"))
    ap.add_argument("--level", type=int, choices=range(4), default=2)
    args = ap.parse_args()
    args.output.mkdir(parents=True, exist_ok=True)
    for kind in ("reflection", "rigid", "partial", "single"):
        path = args.output / f"{kind}.off"
        write_off(path, *demo(kind, args.level))
        print(path)


if __name__ == "__main__":
    main()
