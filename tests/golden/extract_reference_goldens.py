"""Write the reference's exact coboundary known answers as a JSON fixture.

Reads the reference's own test file tests/exterior/derivative.cpp as text
(it is never compiled or run) and extracts, for every test_derivative<...>
call, the grid size, the layout flag, the input/output form degrees, the
dimension, and every EXPECT_EQ(derivative(node, component), value) that
follows it.  The fixture holds only data: inputs are implied by the test
harness rule (component i holds i+1 everywhere except i+2 at node (1,...,1)),
outputs are the expected values.

Usage: python tests/golden/extract_reference_goldens.py [reference_root]
"""
import json
import re
import sys
from pathlib import Path

AXES_BY_DIM = {1: "X", 2: "XY", 3: "XYZ", 4: "TXYZ"}


def index_rank(text):
    text = text.strip()
    m = re.match(r"sil::tensor::TensorAntisymmetricIndex<(.*)>$", text, re.S)
    if m:
        inner = m.group(1).strip()
        return 0 if inner == "" else len([t for t in inner.split(",") if t.strip()])
    if re.match(r"sil::tensor::TensorNaturalIndex<\s*>$", text):
        return 0
    return 1  # a bare natural index such as Mu2


def split_template_args(s):
    args, depth, cur = [], 0, ""
    for ch in s:
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        if ch == "," and depth == 0:
            args.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        args.append(cur)
    return [a.strip() for a in args]


def main():
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    src = (root / "tests/exterior/derivative.cpp").read_text()
    calls = list(re.finditer(r"auto \[(\w+), (\w+)\]\s*=\s*test_derivative<(.*?)>\(\);", src, re.S))
    cases = []
    for ci, call in enumerate(calls):
        args = split_template_args(call.group(3))
        n_pts = int(args[0])
        coalescent = args[1] == "true"
        in_rank = index_rank(args[2])
        out_rank = index_rank(args[3])
        ndim = len(args) - 4
        name = call.group(2)
        end = calls[ci + 1].start() if ci + 1 < len(calls) else len(src)
        body = src[call.end():end]
        test_name = re.findall(r"TEST\(ExteriorDerivative, (\w+)\)", src[: call.start()])[-1]
        expects = []
        pat = re.compile(
            name + r"\(\s*ddc::DiscreteElement<[^>]*>\s*\{([^}]*)\},\s*"
            + name + r"\.accessor\(\)\.access_element<([^>]*)>\(\)\),\s*(-?[0-9.]+)\)",
            re.S,
        )
        axes = AXES_BY_DIM[ndim]
        for m in pat.finditer(body):
            node = [int(v) for v in m.group(1).split(",")]
            comp_axes = [a.strip() for a in m.group(2).split(",")]
            mask = 0
            for a in comp_axes:
                mask |= 1 << axes.index(a)
            expects.append({"node": node, "axes_mask": mask, "value": float(m.group(3))})
        cases.append(
            {
                "test": test_name,
                "points_per_axis": n_pts,
                "index_first": coalescent,
                "in_degree": in_rank,
                "out_degree": out_rank,
                "ndim": ndim,
                "expects": expects,
            }
        )
    out = Path(__file__).with_name("derivative_known_answers.json")
    out.write_text(json.dumps({"source": "tests/exterior/derivative.cpp", "cases": cases}, indent=1))
    print(f"wrote {out} with {len(cases)} cases, {sum(len(c['expects']) for c in cases)} expectations")


if __name__ == "__main__":
    main()
