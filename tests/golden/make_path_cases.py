"""Extract the known-answer cases of the reference's path geometry tests as a fixture (run where the reference
checkout is; the output is committed, the reference never travels).

    python tests/golden/make_path_cases.py /root/reference

Reads src/OpenParEMCommon/path.cpp as text:
  * test_is_point_on_line (path.cpp:236-429): segment end points, the test point, the tolerance and whether the
    point is expected on the segment (the variables are tracked line by line, as the function assigns them);
  * Path::test_is_point_inside_m / _mr / _sqr2 (path.cpp:1357-1756): the closed path each assumes (from the comment
    above it) and the expected inside/outside answer for each test point.
Writes tests/golden/path_cases.json.
"""
import json
import os
import re
import sys

NUM = r"[-+0-9.e]+"


def arith(expr, env):
    """Value of a C expression made of numbers, + - * / and the tracked variables."""
    tokens = re.findall(r"[A-Za-z_][A-Za-z_0-9]*|[0-9.]+(?:e[-+]?[0-9]+)?|[-+*/()]", expr.replace(" ", ""))
    out = []
    for t in tokens:
        if re.match(r"[A-Za-z_]", t):
            out.append(repr(float(env[t])))
        else:
            out.append(t)
    return float(eval("".join(out), {"__builtins__": {}}, {}))  # digits and operators only


def on_line_cases(src):
    start = src.index("void test_is_point_on_line ()")
    body = src[start:src.index("\n}\n", start)]
    env, cases = {}, []
    for line in body.split("\n"):
        line = line.strip()
        call = re.search(r"if \((!?)test_is_point_on_line \((.*)\)\)\s+PetscPrintf", line)
        if call:
            args = [arith(a, env) for a in call.group(2).split(",")]
            expected = call.group(1) != "!"
            if len(args) == 7:   # 2D: xt, yt, x1, y1, x2, y2, tol
                p, a, b, tol = args[0:2], args[2:4], args[4:6], args[6]
            else:                # 3D: xt, yt, zt, x1, y1, z1, x2, y2, z2, tol
                p, a, b, tol = args[0:3], args[3:6], args[6:9], args[9]
            cases.append(dict(p=p, a=a, b=b, tol=tol, on=expected))
            continue
        for name, val in re.findall(r"([a-z][a-z0-9]*)=([^;]+);", line):
            env[name] = arith(val, env)
    return cases


def inside_cases(src, fname):
    start = src.index(f"void Path::{fname} ()")
    head = src[:start]
    comment = head[head.rindex("// Assumes the path:"):]
    pts = [[float(v) for v in m.split(",")] for m in re.findall(r"point=\(([^)]*)\)", comment)]
    closed = "closed=true" in comment
    body = src[start:src.index("\n}\n", start)]
    cases = []
    for xt, yt, zt, exp in re.findall(rf"xt=({NUM}); yt=({NUM}); zt=({NUM}); expected=\"(inside|outside)\"", body):
        cases.append(dict(p=[float(xt), float(yt), float(zt)], inside=exp == "inside"))
    return dict(path=pts, closed=closed, cases=cases)


def main(ref):
    src = open(os.path.join(ref, "src", "OpenParEMCommon", "path.cpp")).read()
    out = dict(source="src/OpenParEMCommon/path.cpp", path_tolerance=1e-11,
               on_line=on_line_cases(src),
               inside={k: inside_cases(src, f"test_is_point_inside_{k}") for k in ("m", "mr", "sqr2")})
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "path_cases.json")
    json.dump(out, open(dst, "w"), indent=1)
    print(dst, len(out["on_line"]), {k: len(v["cases"]) for k, v in out["inside"].items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
