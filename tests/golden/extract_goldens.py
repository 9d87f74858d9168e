"""Extract the reference test suite's golden vectors into ``goldens.json``.

The reference's Boost.Test cases (Tests/poisson.cpp, Tests/input.cpp) hold
MATLAB-computed expected results as literal ``double real_sol[] = {...}``
arrays.  This script reads those test files *as text*, and for every test case
records: the case name, how the matrix is made (Poisson generator, Matrix
Market loader or Kronecker .bin loader, with the exact flags the test passes),
the operation (``mv`` or ``powerIteration`` with its iteration count), which
vector the test checks, the Boost tolerance, the index type, and the expected
values.  Nothing from the reference is executed or imported.

Run it in a checkout that has the reference at /root/reference:

    python tests/golden/extract_goldens.py

The output (``tests/golden/goldens.json``) is committed; tests only read it.
"""
import json
import os
import re
import sys

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "goldens.json")

CASE_RE = re.compile(r"BOOST_AUTO_TEST_CASE\((\w+),\s*\*\s*boost::unit_test::tolerance\(std::pow\(10,\s*(-?\d+)\)\)\)")


def _parse_case(lines, start, relpath):
    m = CASE_RE.search(lines[start])
    name, tol_exp = m.group(1), int(m.group(2))
    case = {"name": name, "tolerance": 10.0 ** tol_exp, "ref": f"{relpath}:{start + 1}"}
    depth = 0
    body = []
    for i in range(start, len(lines)):
        body.append((i, lines[i]))
        depth += lines[i].count("{") - lines[i].count("}")
        if i > start and depth == 0:
            break
    text = "\n".join(l for _, l in body)

    idx = re.search(r"SparseMatrix<double,\s*(int|unsigned int)>", text)
    case["index_type"] = "uint32" if idx and idx.group(1) == "unsigned int" else "int32"

    sol = re.search(r"double real_sol\[\]\s*=\s*\{(.*?)\};", text, re.S)
    vals = [v.strip() for v in sol.group(1).replace("\n", " ").split(",") if v.strip()]
    case["expected"] = [float(v) for v in vals]
    sol_line = next(i for i, l in body if "real_sol[]" in l)
    case["expected_ref"] = f"{relpath}:{sol_line + 1}"

    g = re.search(r"generatePoissonMatrix\((\d+),\s*(\d+),", text)
    mm = re.search(r'loadFromMM\("Test_input/([\w.]+)",\s*(true|false),\s*(true|false)(?:,\s*(true|false))?\)', text)
    bn = re.search(r'loadFromBin\("Test_input/([\w.]+)",\s*(std::pow\(\d+,\s*\d+\)|\d+),\s*(true|false),\s*(true|false)\)', text)
    if g:
        case["matrix"] = {"kind": "poisson", "m": int(g.group(1)), "n": int(g.group(2))}
    elif mm:
        case["matrix"] = {"kind": "mtx", "file": mm.group(1), "has_data": mm.group(2) == "true",
                          "symmetric": mm.group(3) == "true",
                          "random_fill": (mm.group(4) == "true") if mm.group(4) else False}
    elif bn:
        size_expr = bn.group(2).strip()
        pw = re.match(r"std::pow\((\d+),\s*(\d+)\)", size_expr)
        size = int(pw.group(1)) ** int(pw.group(2)) if pw else int(size_expr)
        case["matrix"] = {"kind": "bin", "file": bn.group(1), "size": size,
                          "symmetric": bn.group(3) == "true", "random": bn.group(4) == "true"}
    else:
        raise ValueError(f"no matrix source in {name}")

    p = re.search(r"powerIteration\(x,\s*y,\s*(\d+)\)", text)
    if p:
        case["op"] = "powerIteration"
        case["iterations"] = int(p.group(1))
    else:
        case["op"] = "mv"
    chk = re.search(r"BOOST_TEST\((\w)\[i\]\s*==\s*real_sol\[i\]\)", text)
    case["checks"] = chk.group(1)
    case["start_vector"] = "ones" if "std::fill(x, x+mat_size, 1.)" in text else "unknown"
    return case


def main():
    if not os.path.isdir(REF):
        sys.exit(f"reference checkout not found at {REF}")
    cases = []
    for rel in ("Tests/poisson.cpp", "Tests/input.cpp"):
        with open(os.path.join(REF, rel)) as f:
            lines = f.read().split("\n")
        for i, line in enumerate(lines):
            if CASE_RE.search(line):
                cases.append(_parse_case(lines, i, rel))
    with open(OUT, "w") as f:
        json.dump({"source": "reference Tests/poisson.cpp and Tests/input.cpp (MATLAB golden vectors)",
                   "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
