"""Writes closed_forms.json: inputs and expected outputs of dexel morphology
cases whose answer follows from the definition alone (no oracle involved).

Dilation by a radius r over the integer lattice (VoronoiBruteForce.cpp:39-50):
    D(x, y) = U_{d^2 <= r^2} U_{[a,b] in I(x', y')} [a - sqrt(r^2 - d^2), b + sqrt(r^2 - d^2)]
with d^2 = (x - x')^2 + (y - y')^2, touching intervals fused.

Cases:
  point   -- one zero-length segment [z0, z0]: a lattice ball, column (x, y)
             gets [z0 - h, z0 + h], h = sqrt(r^2 - d^2), for d^2 <= r^2
  capsule -- one segment [a, b]: [a - h, b + h]
  gap     -- two segments in one column at gap g: one interval iff 2r >= g
             at the column itself, columns further out see the gap close later
  box     -- erosion of a box of full columns: a column survives iff its
             whole lattice disk (|dx|, |dy| <= floor(r)) lies in the box, and
             keeps [z0 + r, z1 - r]

    python tests/golden/make_golden.py [out_dir]   # (re)writes out_dir/closed_forms.json
"""

import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def grid_cols(n, fill):
    return [fill(x, y) for y in range(n) for x in range(n)]


def point_case(n, c, z0, r):
    def exp(x, y):
        d2 = (x - c[0]) ** 2 + (y - c[1]) ** 2
        if d2 > r * r:
            return []
        h = math.sqrt(r * r - d2)
        return [z0 - h, z0 + h]
    inp = grid_cols(n, lambda x, y: [z0, z0] if (x, y) == tuple(c) else [])
    return {"name": f"point_r{r:g}", "op": "dilation", "radius": r, "nx": n, "ny": n,
            "z_range": [0.0, 64.0], "input": inp, "expected": grid_cols(n, exp)}


def capsule_case(n, c, a, b, r):
    def exp(x, y):
        d2 = (x - c[0]) ** 2 + (y - c[1]) ** 2
        if d2 > r * r:
            return []
        h = math.sqrt(r * r - d2)
        return [a - h, b + h]
    inp = grid_cols(n, lambda x, y: [a, b] if (x, y) == tuple(c) else [])
    return {"name": f"capsule_r{r:g}", "op": "dilation", "radius": r, "nx": n, "ny": n,
            "z_range": [0.0, 64.0], "input": inp, "expected": grid_cols(n, exp)}


def gap_case(n, c, a, b, g, r):
    # segments [a, b] and [b + g, b + g + (b - a)] in column c
    a2, b2 = b + g, b + g + (b - a)

    def exp(x, y):
        d2 = (x - c[0]) ** 2 + (y - c[1]) ** 2
        if d2 > r * r:
            return []
        h = math.sqrt(r * r - d2)
        if b + h >= a2 - h:  # the gap closes (touching fuses)
            return [a - h, b2 + h]
        return [a - h, b + h, a2 - h, b2 + h]
    inp = grid_cols(n, lambda x, y: [a, b, a2, b2] if (x, y) == tuple(c) else [])
    return {"name": f"gap{g:g}_r{r:g}", "op": "dilation", "radius": r, "nx": n, "ny": n,
            "z_range": [0.0, 64.0], "input": inp, "expected": grid_cols(n, exp)}


def box_case(n, lo, hi, z0, z1, r, zr):
    k = int(math.floor(r))

    def inside(x, y):
        return lo[0] <= x <= hi[0] and lo[1] <= y <= hi[1]

    def exp(x, y):
        if lo[0] + k <= x <= hi[0] - k and lo[1] + k <= y <= hi[1] - k and z0 + r <= z1 - r:
            return [z0 + r, z1 - r]
        return []
    inp = grid_cols(n, lambda x, y: [z0, z1] if inside(x, y) else [])
    return {"name": f"box_erosion_r{r:g}", "op": "erosion", "radius": r, "nx": n, "ny": n,
            "z_range": list(zr), "input": inp, "expected": grid_cols(n, exp)}


def main(out_dir=HERE):
    cases = [
        point_case(15, (7, 6), 20.0, 4.0),
        point_case(15, (3, 9), 31.5, 5.5),
        capsule_case(13, (6, 6), 10.25, 17.75, 3.0),
        capsule_case(13, (0, 12), 5.0, 9.0, 6.2),
        gap_case(17, (8, 8), 10.0, 14.0, 6.0, 3.5),
        gap_case(17, (8, 8), 10.0, 14.0, 6.0, 2.9),
        box_case(20, (4, 3), (15, 16), 12.0, 40.0, 3.0, (0.0, 52.0)),
        box_case(20, (2, 2), (17, 17), 8.0, 44.0, 4.5, (0.0, 52.0)),
    ]
    with open(os.path.join(out_dir, "closed_forms.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": cases}, f, indent=None,
                  separators=(",", ":"))
        f.write("\n")
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else HERE)
