#!/usr/bin/env python3
"""Extract the reference's own golden data into small JSON fixtures.

This script is the only committed file that reads the reference tree.  It
reads two *data* files as text and writes their numbers out as JSON:

* ``ExaHyPE/exahype/tests/kernels/testdata/generic_euler_testdata.cpp`` --
  the golden vectors used by the reference's kernel unit tests
  (``GenericEulerKernelTest2D.cpp`` / ``GenericEulerKernelTest3D.cpp``).
  Every ``const double NAME[N] = {...};`` inside a ``namespace test...`` block
  is kept, grouped by the ``#ifdef Dim2`` / ``#ifdef Dim3`` section it sits in.
* ``ExaHyPE/exahype/tests/kernels/testdata/elasticity_testdata.cpp`` -- the
  2D (order 4, 9 variables + 3 parameters) vectors of the linear-kernel tests
  in ``ElasticityKernelTest2D.cpp`` (namespaces testRiemannSolverLinear,
  testSpaceTimePredictorLinear, testVolumeIntegralLinear,
  testSurfaceIntegralLinear).  The reference runs those tests with
  ``validate(false)``; the vectors still match the generic linear kernels.
* ``ExaHyPE/kernels/generated/GaussLegendreBasis.csnippet`` -- the operator
  tables (nodes, weights, Kxi, iK1, dudx, FCoeff, fineGridProjector,
  equidistantGridProjector) that the reference generator printed with ~64
  significant digits, for orders 0..MAX_ORDER.

Decimal strings are converted with Python's correctly rounded ``float()``,
which is the same round-to-nearest a C compiler applies to the literals.

Usage:  python tests/golden/extract_reference_fixtures.py [/root/reference]
Outputs: tests/golden/reference_testdata.json, tests/golden/legendre_tables.json,
         tests/golden/elasticity_testdata.json
"""
import json
import os
import re
import sys

MAX_ORDER = 9
HERE = os.path.dirname(os.path.abspath(__file__))


def _numbers(body):
    out = []
    body = re.sub(r"//[^\n]*", " ", body)          # line comments
    body = re.sub(r"/\*.*?\*/", " ", body, flags=re.S)  # block comments
    for tok in body.replace("\n", " ").split(","):
        tok = tok.strip()
        if tok:
            out.append(float(tok))
    return out


def extract_testdata(ref):
    path = os.path.join(ref, "ExaHyPE/exahype/tests/kernels/testdata/generic_euler_testdata.cpp")
    text = open(path).read()
    result = {"dim2": {}, "dim3": {}}
    section = None
    namespace = None
    pos = 0
    # Walk the file line by line for #ifdef / namespace bookkeeping, and use a
    # regex for the array literals (which span many lines).
    array_re = re.compile(r"const\s+double\s+(\w+)\s*\[\s*(\d+)\s*\]\s*=?\s*\{(.*?)\}\s*;", re.S)
    tokens = re.compile(r"(#ifdef\s+Dim2|#ifdef\s+Dim3|#endif[^\n]*|namespace\s+(test\w+)\s*\{|"
                        r"const\s+double\s+\w+\s*\[\s*\d+\s*\]\s*=?\s*\{.*?\}\s*;)", re.S)
    for m in tokens.finditer(text):
        tok = m.group(1)
        if tok.startswith("#ifdef"):
            section = "dim2" if tok.endswith("Dim2") else "dim3"
        elif tok.startswith("#endif"):
            section = None
        elif tok.startswith("namespace"):
            namespace = m.group(2)
        else:
            a = array_re.match(tok)
            name, n, body = a.group(1), int(a.group(2)), a.group(3)
            vals = _numbers(body)
            if len(vals) != n:
                raise RuntimeError(f"{namespace}::{name}: expected {n} values, got {len(vals)}")
            if section is None:
                raise RuntimeError(f"array {name} outside a Dim section")
            result[section].setdefault(namespace, {})[name] = vals
    return result


def extract_elasticity(ref):
    path = os.path.join(ref, "ExaHyPE/exahype/tests/kernels/testdata/elasticity_testdata.cpp")
    text = open(path).read()
    array_re = re.compile(r"const\s+double\s+(\w+)\s*\[([^\]]*)\]\s*=?\s*\{(.*?)\}\s*;", re.S)
    tokens = re.compile(r"(namespace\s+(test\w+)\s*\{|const\s+double\s+\w+\s*\[[^\]]*\]\s*=?\s*\{.*?\}\s*;)", re.S)
    result = {}
    namespace = None
    for m in tokens.finditer(text):
        tok = m.group(1)
        if tok.startswith("namespace"):
            namespace = m.group(2)
            continue
        a = array_re.match(tok)
        name, size, body = a.group(1), a.group(2), a.group(3)
        n = 1
        for f in size.split("*"):
            n *= int(f)
        vals = _numbers(body)
        if len(vals) != n:
            raise RuntimeError(f"{namespace}::{name}: expected {n} values, got {len(vals)}")
        result.setdefault(namespace, {})[name] = vals
    return result


def extract_tables(ref):
    path = os.path.join(ref, "ExaHyPE/kernels/generated/GaussLegendreBasis.csnippet")
    text = open(path).read()
    arrays = {}
    for m in re.finditer(r"double\s+kernels::legendre::(\w+)\s*\[[^\]]*\]\s*=\s*\{([^}]*)\}\s*;", text):
        name, body = m.group(1), m.group(2)
        arrays[name] = _numbers(body)
    tables = {}
    for order in range(MAX_ORDER + 1):
        b = order + 1
        t = {
            "weights": arrays[f"weights_{order}"],
            "nodes": arrays[f"nodes_{order}"],
            "Kxi": [arrays[f"Kxi_{order}_{i}"] for i in range(b)],
            "iK1": [arrays[f"iK1_{order}_{i}"] for i in range(b)],
            "dudx": [arrays[f"dudx_{order}_{i}"] for i in range(b)],
            "equidistantGridProjector": [arrays[f"equidistantGridProjector_{order}_{i}"] for i in range(b)],
            "FCoeff": [arrays[f"FCoeff_0_{order}"], arrays[f"FCoeff_1_{order}"]],
            "fineGridProjector": [[arrays[f"fineGridProjector_{j}_{order}_{i}"] for i in range(b)]
                                  for j in range(3)],
        }
        for k in ("Kxi", "iK1", "dudx", "fineGridProjector"):
            rows = t[k] if k != "fineGridProjector" else [r for sub in t[k] for r in sub]
            assert all(len(r) == b for r in rows), (order, k)
        tables[str(order)] = t
    return tables


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    td = extract_testdata(ref)
    with open(os.path.join(HERE, "reference_testdata.json"), "w") as f:
        json.dump({"source": "ExaHyPE/exahype/tests/kernels/testdata/generic_euler_testdata.cpp",
                   "data": td}, f, separators=(",", ":"))
    el = extract_elasticity(ref)
    with open(os.path.join(HERE, "elasticity_testdata.json"), "w") as f:
        json.dump({"source": "ExaHyPE/exahype/tests/kernels/testdata/elasticity_testdata.cpp",
                   "data": el}, f, separators=(",", ":"))
    print("elasticity", {k: {n: len(v) for n, v in arrs.items()} for k, arrs in el.items()})
    tb = extract_tables(ref)
    with open(os.path.join(HERE, "legendre_tables.json"), "w") as f:
        json.dump({"source": "ExaHyPE/kernels/generated/GaussLegendreBasis.csnippet",
                   "tables": tb}, f, separators=(",", ":"))
    for sec, d in td.items():
        print(sec, {k: {n: len(v) for n, v in arrs.items()} for k, arrs in d.items()})
    print("tables: orders", sorted(int(k) for k in tb))


if __name__ == "__main__":
    main()
