"""Transcribe the reference's matmul golden vectors into JSON fixtures.

Reads the reference's JUnit test sources *as text* (it never compiles or runs
them) and writes the literal inputs and expected outputs of each hot-path test
into ``tests/golden/matmul_golden.json``.  Numbers are kept as the decimal
strings written in the tests; the consumer parses them with Python's
``float()``, which is correctly rounded, as is Java's ``Double.parseDouble``
(used by ``Complex128(String)`` via ComplexNumberParser.java:112-158).

Usage (in a container that has the reference checked out):

    python tests/golden/transcribe_fixtures.py /root/reference

Test files read (T: = src/test/java/org/flag4j/):
  T:linalg/ops/dense/real/MatrixMultiplyTests.java:14-135
  T:arrays/dense/matrix/MatrixMultTests.java:24-63, :193-246, :250-283
  T:linalg/ops/dense/real/RealDenseMatMultTransposeTests.java:13-69
  T:linalg/ops/dense/complex/CMatrixMultiplyTests.java:17-169
  T:arrays/dense/complex_matrix/CMatrixMultTests.java:64-102, :198-256
  T:linalg/ops/dense/complex/ComplexDenseMatMultTransposeTests.java:40-107
  T:arrays/dense/matrix/MatrixMultTests.java:66-102 (Matrix.mult(CMatrix))
  T:arrays/dense/complex_matrix/CMatrixMultTests.java:20-62 (CMatrix.mult(Matrix))
  T:arrays/dense/matrix/MatrixVectorTests.java:27-98 (Matrix.mult(Vector), Matrix.mult(CVector))
  T:arrays/dense/complex_matrix/CMatrixMatVecMultTests.java:22-100 (CMatrix.mult(Vector / CVector))
  T:linalg/solvers/RealLstsqSolverTests.java:17-98 (RealLstsqSolver.solve, vector and matrix right-hand sides)
"""
import json
import os
import re
import sys

T = "src/test/java/org/flag4j/"

# (fixture name, file, method, {field: (java variable, occurrence within method)})
SPECS = [
    ("real_square", "linalg/ops/dense/real/MatrixMultiplyTests.java", "squareTestCase",
     {"a": ("entriesA", 0), "b": ("entriesB", 0), "c": ("exp", 0)}),
    ("real_rect", "linalg/ops/dense/real/MatrixMultiplyTests.java", "rectangleTestCase",
     {"a": ("entriesA", 0), "b": ("entriesB", 0), "c": ("exp", 0)}),
    ("real_colvec", "linalg/ops/dense/real/MatrixMultiplyTests.java", "columnVectorTestCase",
     {"a": ("entriesA", 0), "b": ("entriesB", 0), "c": ("exp", 0)}),
    ("real_api_mult", "arrays/dense/matrix/MatrixMultTests.java", "matMultTestCase",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "c": ("expEntries", 0),
      "a_bad": ("aEntries", 1), "b_bad": ("bEntries", 1)}),
    ("real_api_pow", "arrays/dense/matrix/MatrixMultTests.java", "powTestCase",
     {"a": ("aEntries", 0), "c3": ("expEntries", 0), "c0": ("expEntries", 1),
      "c1": ("expEntries", 2), "a_nonsquare": ("aEntries", 3)}),
    ("real_api_multT", "arrays/dense/matrix/MatrixMultTests.java", "multTransposeTestCase",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "a_bad": ("aEntries", 1), "b_bad": ("bEntries", 1)}),
    ("real_T_square", "linalg/ops/dense/real/RealDenseMatMultTransposeTests.java", "squareTestCase",
     {"a": ("entriesA", 0), "b": ("entriesB", 0)}),
    ("real_T_rect", "linalg/ops/dense/real/RealDenseMatMultTransposeTests.java", "rectangleTestCase",
     {"a": ("entriesA", 0), "b": ("entriesB", 0)}),
    ("cplx_square", "linalg/ops/dense/complex/CMatrixMultiplyTests.java", "squareTestCase",
     {"a": ("entriesA", 0), "b": ("entriesB", 0), "c": ("exp", 0)}),
    ("cplx_rect", "linalg/ops/dense/complex/CMatrixMultiplyTests.java", "rectangleTestCase",
     {"a": ("entriesA", 0), "b": ("entriesB", 0), "c": ("exp", 0)}),
    ("cplx_colvec", "linalg/ops/dense/complex/CMatrixMultiplyTests.java", "columnVectorTestCase",
     {"a": ("entriesA", 0), "b": ("entriesB", 0), "c": ("exp", 0)}),
    ("cplx_api_mult", "arrays/dense/complex_matrix/CMatrixMultTests.java", "matMultComplexTestCase",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "c": ("expEntries", 0),
      "a_bad": ("aEntries", 1), "b_bad": ("bEntries", 1)}),
    ("cplx_api_pow", "arrays/dense/complex_matrix/CMatrixMultTests.java", "powTestCase",
     {"a": ("aEntries", 0), "c3": ("expEntries", 0), "c0": ("expEntries", 1),
      "c1": ("expEntries", 2), "a_nonsquare": ("aEntries", 3)}),
    ("cplx_T_square", "linalg/ops/dense/complex/ComplexDenseMatMultTransposeTests.java", "squareTestCase",
     {"a": ("entriesA", 0), "b": ("entriesB", 0)}),
    ("cplx_T_rect", "linalg/ops/dense/complex/ComplexDenseMatMultTransposeTests.java", "rectangleTestCase",
     {"a": ("entriesA", 0), "b": ("entriesB", 0)}),
    ("real_cmat_api_mult", "arrays/dense/matrix/MatrixMultTests.java", "matMultComplexTestCase",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "c": ("expCEntries", 0), "b_bad": ("bEntries", 1)}),
    ("cplx_rmat_api_mult", "arrays/dense/complex_matrix/CMatrixMultTests.java", "matMultTestCase",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "c": ("expEntries", 0), "b_bad": ("bEntries", 1)}),
    ("real_api_matvec", "arrays/dense/matrix/MatrixVectorTests.java", "matVecMultTestCase",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "c": ("expEntries", 0), "b_bad": ("bEntries", 1)}),
    ("real_cvec_api_matvec", "arrays/dense/matrix/MatrixVectorTests.java", "matVecMultComplexTestCase",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "c": ("expCEntries", 0), "b_bad": ("bEntries", 1)}),
    ("cplx_vec_api_matvec", "arrays/dense/complex_matrix/CMatrixMatVecMultTests.java", "matMultTestCase",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "c": ("expEntries", 0), "b_bad": ("bEntries", 1)}),
    ("cplx_api_matvec", "arrays/dense/complex_matrix/CMatrixMatVecMultTests.java", "matMultComplexTestCase",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "c": ("expEntries", 0), "b_bad": ("bEntries", 1)}),
    ("lstsq_vec_1", "linalg/solvers/RealLstsqSolverTests.java", "solveTestCase",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "x": ("expEntries", 0)}),
    ("lstsq_vec_2", "linalg/solvers/RealLstsqSolverTests.java", "solveTestCase",
     {"a": ("aEntries", 1), "b": ("bEntries", 1), "x": ("expEntries", 1)}),
    ("lstsq_mat_1", "linalg/solvers/RealLstsqSolverTests.java", "solveTestMatrix",
     {"a": ("aEntries", 0), "b": ("bEntries", 0), "x": ("expEntries", 0)}),
    ("lstsq_mat_2", "linalg/solvers/RealLstsqSolverTests.java", "solveTestMatrix",
     {"a": ("aEntries", 1), "b": ("bEntries", 1), "x": ("expEntries", 1)}),
]


def method_body(src, name):
    m = re.search(r"void\s+" + name + r"\s*\(\s*\)\s*\{", src)
    if not m:
        raise KeyError(name)
    i, depth = m.end(), 1
    while depth:
        depth += {"{": 1, "}": -1}.get(src[i], 0)
        i += 1
    return src[m.end():i - 1], src[:m.start()].count("\n") + 1


def literal_after(body, var, occ):
    hits = [m for m in re.finditer(r"\b" + var + r"\s*=\s*new\s+\w+(\[\])+\s*\{", body)]
    m = hits[occ]
    i, depth = m.end(), 1
    while depth:
        depth += {"{": 1, "}": -1}.get(body[i], 0)
        i += 1
    return body[m.end() - 1:i], body[:m.start()].count("\n")


def split_top(s):
    """Split a brace-free comma list at depth 0 of parentheses."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def num(tok):
    tok = tok.strip()
    if tok == "Math.PI":
        return "3.141592653589793"   # Math.PI is the double nearest pi
    float(tok)
    return tok


def parse_complex_string(s):
    """Restates ComplexNumberParser.java:112-158 for the literal forms the tests use."""
    s = s.replace(" ", "")
    if s.endswith("i"):
        body = s[:-1]
        # split at the last +/- that is not a leading sign or exponent sign
        cut = None
        for k in range(len(body) - 1, 0, -1):
            if body[k] in "+-" and body[k - 1] not in "eE":
                cut = k
                break
        if cut is None:
            re_s, im_s = "0.0", body if body not in ("", "+", "-") else body + "1"
        else:
            re_s, im_s = body[:cut], body[cut:]
            if im_s in ("+", "-"):
                im_s += "1"
        return [num(re_s), num(im_s)]
    return [num(s), "0.0"]


def parse_scalar(tok):
    tok = tok.strip()
    if tok == "Complex128.ZERO":
        return ["0.0", "0.0"]
    m = re.fullmatch(r"new\s+Complex128\((.*)\)", tok, re.S)
    if m:
        args = split_top(m.group(1))
        if len(args) == 1 and args[0].startswith('"'):
            return parse_complex_string(args[0].strip('"'))
        if len(args) == 1:
            return [num(args[0]), "0.0"]
        return [num(args[0]), num(args[1])]
    return num(tok)


def parse_literal(lit):
    lit = lit.strip()
    assert lit[0] == "{" and lit[-1] == "}"
    inner = lit[1:-1].strip()
    if inner.startswith("{"):
        rows, depth, start = [], 0, None
        for i, ch in enumerate(inner):
            if ch == "{":
                if depth == 0:
                    start = i
                depth += 1
            elif ch == "}":
                depth -= 1
                if depth == 0:
                    rows.append(parse_literal(inner[start:i + 1]))
        return rows
    return [parse_scalar(t) for t in split_top(inner)]


def main(ref_root):
    out = {}
    for name, rel, meth, fields in SPECS:
        path = os.path.join(ref_root, T, rel)
        with open(path) as f:
            src = f.read()
        body, line0 = method_body(src, meth)
        entry = {"source": "T:" + rel + ":" + meth}
        for field, (var, occ) in fields.items():
            lit, off = literal_after(body, var, occ)
            entry[field] = parse_literal(lit)
            entry[field + "_line"] = line0 + off
        out[name] = entry
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "matmul_golden.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst, "with", len(out), "fixtures")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
