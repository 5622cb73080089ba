"""Writes the golden fixtures under tests/golden/ from the reference's own test
outputs (data files only; run in a container where the IBAMR checkout is
mounted at /root/reference).

  interpolate_01_{2d,3d}.{ib_4,ib_6,bspline_3}.json
      from tests/interpolate/interpolate_01_<dim>.<kernel>.output: the printed
      point coordinates and interpolated values (cell-centred, depth NDIM).
  ghost_accumulation_01_3d.side.spread.a.json
      from tests/IBTK/ghost_accumulation_01_3d.side.spread.a.output: the side
      data values over the patch box after a PIECEWISE_LINEAR spread.

Usage: python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def interp_fixture(ref, dim, kernel):
    path = os.path.join(ref, "tests", "interpolate", "interpolate_01_%s.%s.output" % (dim, kernel))
    rows = []
    with open(path) as f:
        header = f.readline()
        assert header.startswith("x, y, z")
        for line in f:
            line = line.strip()
            if line:
                rows.append([float(t) for t in line.split(",")])
    return {"source": "tests/interpolate/interpolate_01_%s.%s.output" % (dim, kernel),
            "columns": ["x", "y", "z", "u0", "u1", "u2"], "rows": rows}


def side_fixture(ref, name):
    path = os.path.join(ref, "tests", "IBTK", name + ".output")
    axes = {}
    cur = None
    pat = re.compile(r"array\(([-\d,]+)\) = (\S+)")
    with open(path) as f:
        for line in f:
            m = re.match(r"Array side normal = (\d+)", line)
            if m:
                cur = int(m.group(1))
                axes[cur] = []
                continue
            m = pat.search(line)
            if m and cur is not None:
                idx = [int(t) for t in m.group(1).split(",")]
                axes[cur].append(idx + [float(m.group(2))])
    return {"source": "tests/IBTK/%s.output" % name, "axes": {str(k): v for k, v in axes.items()}}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for dim in ("2d", "3d"):
        for kernel in ("ib_4", "ib_6", "bspline_3"):
            fx = interp_fixture(ref, dim, kernel)
            with open(os.path.join(HERE, "interpolate_01_%s.%s.json" % (dim, kernel)), "w") as f:
                json.dump(fx, f)
    for name in ("ghost_accumulation_01_3d.side.spread.a", "ghost_accumulation_01_2d.side.spread.a"):
        fx = side_fixture(ref, name)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(fx, f)


if __name__ == "__main__":
    main()
