"""Writes the golden fixtures under tests/golden/ from the reference's own test
outputs (data files only; run in a container where the IBAMR checkout is
mounted at /root/reference).

  interpolate_01_{2d,3d}.{ib_4,ib_6,bspline_3,bspline_4,bspline_5,bspline_6}.json
      from tests/interpolate/interpolate_01_<dim>.<kernel>.output: the printed
      point coordinates and interpolated values (cell-centred, depth NDIM).
  curve2d_64.vertex, curve2d_64.spring
      copied from tests/IB/ (IBStandardInitializer input files: a closed curve
      of 304 nodes joined by 304 springs).
  ghost_accumulation_01_<case>.patches.json (8 cases, 2-D/3-D, cell/side,
      one or several patches, periodic or not): every patch of the finest
      level after the ghost accumulation, with the input settings it used.
  ghost_accumulation_01_3d.side.spread.a.json
      from tests/IBTK/ghost_accumulation_01_3d.side.spread.a.output: the side
      data values over the patch box after a PIECEWISE_LINEAR spread.

Usage: python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import re
import shutil
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


CASES = ["2d.cell.spread.a", "2d.cell.spread.b", "2d.side.spread.a", "2d.side.spread.b", "3d.cell.spread.a",
         "3d.cell.spread.b", "3d.side.spread.a", "3d.side.spread.b"]


def _input_params(path):
    """The few input-file settings the test depends on (as data)."""
    txt = open(path).read()
    txt = re.sub(r"//[^\n]*", "", txt)
    def val(key):
        m = re.search(r"^\s*%s\s*=\s*(.+)$" % re.escape(key), txt, flags=re.M)
        return m.group(1).strip() if m else None
    out = {"N": int(val("N")), "ghost_width": int(val("ghost_width")), "depth": int(val("depth")),
           "var_type": val("var_type").strip('"'), "kernel": val("IB_DELTA_FUNCTION").strip('"'),
           "periodic": [int(t) for t in val("periodic_dimension").split(",")],
           "max_levels": int(val("max_levels"))}
    m = re.search(r"level_1\s*=\s*([\d,\s]+)", txt)
    out["ratio"] = int(m.group(1).split(",")[0]) if (m and out["max_levels"] > 1) else 1
    return out


def patches_fixture(ref, case):
    """All patches of the finest level printed by ghost_accumulation_01 (after
    SAMRAIGhostDataAccumulator::accumulateGhostData): x_lower, x_upper and the
    values over the patch box, per side axis (side data) or depth (cell data)."""
    name = "ghost_accumulation_01_" + case
    params = _input_params(os.path.join(ref, "tests", "IBTK", name + ".input"))
    lines = open(os.path.join(ref, "tests", "IBTK", name + ".output")).read().splitlines()
    ndim = 2 if case.startswith("2d") else 3
    patches = []
    pat = re.compile(r"array\(([-\d,]+)\) = (\S+)")
    i = 0
    cur = None
    while i < len(lines):
        line = lines[i].strip()
        if line.startswith("Rank:"):
            cur = {"x_lower": [], "x_upper": [], "arrays": {}}
            patches.append(cur)
        elif line == "x lower:":
            cur["x_lower"] = [float(lines[i + 1 + d]) for d in range(ndim)]
            i += ndim
        elif line == "x upper:":
            cur["x_upper"] = [float(lines[i + 1 + d]) for d in range(ndim)]
            i += ndim
        else:
            m = re.match(r"Array (side normal|depth) = (\d+)", line)
            if m and (m.group(1) == "side normal" or params["var_type"] == "CELL"):
                # side data: one array per side normal (depth 1); cell data: per depth
                cur_key = m.group(2)
                cur["arrays"].setdefault(cur_key, [])
            else:
                m = pat.search(line)
                if m and cur is not None:
                    cur["arrays"][cur_key].append([int(t) for t in m.group(1).split(",")] + [float(m.group(2))])
        i += 1
    return {"source": "tests/IBTK/%s.output (+ .input settings)" % name, "ndim": ndim, "params": params,
            "patches": patches}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for dim in ("2d", "3d"):
        for kernel in ("ib_4", "ib_6", "bspline_3", "bspline_4", "bspline_5", "bspline_6"):
            fx = interp_fixture(ref, dim, kernel)
            with open(os.path.join(HERE, "interpolate_01_%s.%s.json" % (dim, kernel)), "w") as f:
                json.dump(fx, f)
    # structure input files of the reference's IB tests (data, copied as they are)
    for fn in ("curve2d_64.vertex", "curve2d_64.spring"):
        shutil.copyfile(os.path.join(ref, "tests", "IB", fn), os.path.join(HERE, fn))
    for case in CASES:
        fx = patches_fixture(ref, case)
        with open(os.path.join(HERE, "ghost_accumulation_01_%s.patches.json" % case), "w") as f:
            json.dump(fx, f)
    for name in ("ghost_accumulation_01_3d.side.spread.a", "ghost_accumulation_01_2d.side.spread.a"):
        fx = side_fixture(ref, name)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(fx, f)


if __name__ == "__main__":
    main()
