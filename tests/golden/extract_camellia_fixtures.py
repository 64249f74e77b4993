"""Writes the golden vectors that the reference's own tests hold, as JSON data.

Run in the build container (it reads the reference sources as text):
    python tests/golden/extract_camellia_fixtures.py /root/reference

Outputs (committed):
  analytic_boundary_integral_cubic.json
      drivers/DPGTests/TestBilinearFormAnalyticBoundaryIntegral.h:83-1673
      (Mathematica values for b(uhat, v) = <uhat n, grad v> on [-1,1]^2, cubic
      SPECTRAL line traces and quad test, MathInnerProduct (v,v)+(grad v, grad v);
      used by DPGTests.cpp:966-1098 with tol 3.76e-14). Arrays are the
      FieldContainers with the cell index dropped:
        pre_stiffness[test][trial], ip_matrix[test][test],
        opt_test_weights[trial][test], final_stiffness[trial][trial];
      each for "nonconforming" traces (16 trial dofs) and, where present,
      "conforming" traces (12 trial dofs).
  factored_cholesky_kats.json
      unit_tests/BFTests.cpp:28-72,136-381 (hard-coded G, B, l with their
      expected K = B^T G^-1 B and rhs = B^T G^-1 l, tol 1e-14).
  math_inner_product_c1.json
      drivers/DPGTests/DPGTests.cpp:755-813 (C1 quad, vertex-ordered basis,
      MathInnerProduct matrix 10/9, 1/18, -2/9).
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_fieldcontainer_fixture(path):
    func_re = re.compile(r"static void (\w+)\(")
    assign_re = re.compile(r"(\w+)\(0,(\d+),(\d+)\)\s*=\s*([-+0-9.eE]+);")
    out = {}
    func = None
    branch = "only"
    for line in open(path):
        m = func_re.search(line)
        if m:
            func = m.group(1)
            branch = "only"
            continue
        if "if (! conformingTraces)" in line:
            branch = "nonconforming"
            continue
        if func and branch == "nonconforming" and line.strip().startswith("else"):
            branch = "conforming"
            continue
        m = assign_re.search(line)
        if m and func:
            i, j, v = int(m.group(2)), int(m.group(3)), float(m.group(4))
            out.setdefault(func, {}).setdefault(branch, {})[(i, j)] = v
    return out


def to_dense(entries):
    ni = max(i for i, _ in entries) + 1
    nj = max(j for _, j in entries) + 1
    a = [[0.0] * nj for _ in range(ni)]
    for (i, j), v in entries.items():
        a[i][j] = v
    return a


def main(ref):
    src = os.path.join(ref, "drivers/DPGTests/TestBilinearFormAnalyticBoundaryIntegral.h")
    raw = parse_fieldcontainer_fixture(src)
    names = {
        "expectedPreStiffnessForCubicsOnQuad": "pre_stiffness",
        "expectedIPMatrixForCubicsOnQuad": "ip_matrix",
        "expectedOptTestWeightsForCubicsOnQuad": "opt_test_weights",
        "expectedFinalStiffnessForCubicsOnQuad": "final_stiffness",
    }
    fx = {"source": "drivers/DPGTests/TestBilinearFormAnalyticBoundaryIntegral.h", "tol": 3.76e-14}
    for fn, key in names.items():
        for branch, entries in raw[fn].items():
            fx.setdefault(key, {})[branch] = to_dense(entries)
    with open(os.path.join(HERE, "analytic_boundary_integral_cubic.json"), "w") as f:
        json.dump(fx, f, indent=1)

    # BFTests.cpp KATs: G and B in the column-major order the test writes them
    kats = {
        "source": "unit_tests/BFTests.cpp:28-72,136-381",
        "tol": 1e-14,
        "cases": [
            {"name": "Identities", "N": 5, "M": 4,
             "G_colmajor": [1.0 if i == j else 0.0 for j in range(5) for i in range(5)],
             "B_colmajor": [1.0 if i == j else 0.0 for j in range(4) for i in range(5)],
             "l": [1.0] * 5,
             "K": [[1.0 if i == j else 0.0 for j in range(4)] for i in range(4)],
             "rhs": [1.0] * 4},
            {"name": "SimpleRectangularMatrices", "N": 3, "M": 2,
             "G_colmajor": [1.0, 0.5, 0.0, 0.5, 1.0, 2.0, 0.0, 2.0, 8.0],
             "B_colmajor": [3, 1, 2, 2, 0, 2], "l": [1, 2, 3],
             "K": [[13.5, 12.5], [12.5, 13.5]], "rhs": [0.75, -1.75]},
            {"name": "SimpleSquareMatrices", "N": 2, "M": 2,
             "G_colmajor": [1.0, 0.5, 0.5, 1.0], "B_colmajor": [3, 1, 2, 0], "l": [1, 2],
             "K": [[28.0 / 3.0, 20.0 / 3.0], [20.0 / 3.0, 16.0 / 3.0]], "rhs": [2.0, 0.0]},
            {"name": "SimpleSquareMatrices2", "N": 2, "M": 2,
             "G_colmajor": [2.0, 0.0, 0.0, 4.0], "B_colmajor": [3, 1, 2, 0], "l": [1, 2],
             "K": [[4.75, 3.0], [3.0, 2.0]], "rhs": [2.0, 1.0]},
        ],
    }
    with open(os.path.join(HERE, "factored_cholesky_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)

    # DPGTests.cpp:755-813: C1 quad (vertex order), (v,v)+(grad v, grad v)
    e = {}
    for i in range(4):
        e[(i, i)] = 10.0 / 9.0
    for (i, j) in [(0, 3), (1, 2), (2, 1), (3, 0), (0, 1), (1, 0), (2, 3), (3, 2)]:
        e[(i, j)] = 1.0 / 18.0
    for (i, j) in [(0, 2), (1, 3), (2, 0), (3, 1)]:
        e[(i, j)] = -2.0 / 9.0
    with open(os.path.join(HERE, "math_inner_product_c1.json"), "w") as f:
        json.dump({"source": "drivers/DPGTests/DPGTests.cpp:755-813", "tol": 1e-14,
                   "vertex_order": [[-1, -1], [1, -1], [1, 1], [-1, 1]],
                   "ip_matrix": to_dense(e)}, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
