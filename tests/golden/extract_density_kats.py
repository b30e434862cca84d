"""Extracts the known-answer cases of the reference's section-list tests
(unittests/marching_steps.cpp:68-938, TEMPLATE_TEST_CASE "dgmr marching step
DensityArray") into density_kats.json: for each SECTION, the list capacity and
start, the sequence of insertions, and every expected size / field value, as
data.  Literals keep their suffix so a float32 literal (e.g. 0.2f) can be
rounded the way the C++ test rounds it.

Run in the build container (the reference is not on the GPU host):
    python tests/golden/extract_density_kats.py /root/reference/unittests/marching_steps.cpp
"""
import json
import os
import re
import sys

NUM = r"-?[0-9]*\.?[0-9]+(?:e-?[0-9]+)?f?"


def lit(tok: str) -> dict:
    tok = tok.strip()
    return {"value": float(tok.rstrip("f")), "f32": tok.endswith("f")}


def main(path: str, out: str) -> None:
    lines = open(path).read().splitlines()
    start, stop = 67, 938  # 0-based slice of the template test case
    cases = []
    section = None
    case = None
    for ln in lines[start:stop]:
        s = re.sub(r"scalar_t\((" + NUM + r")\)", r"\1", ln.strip())
        m = re.match(r'SECTION\("(.*)"\)', s)
        if m:
            section = m.group(1)
            case = None
            continue
        m = re.match(r"dgmr::marching_steps::DensityArray<(\d+), scalar_t> arr\((" + NUM + r")\);", s)
        if m:
            case = {"section": section, "kind": "list", "cap": int(m.group(1)), "start": lit(m.group(2)), "ops": []}
            cases.append(case)
            continue
        m = re.match(r"arr\.put\(\{\s*(" + NUM + r"),\s*(" + NUM + r"),\s*(" + NUM + r")\s*\}\);", s)
        if m:
            case["ops"].append({"op": "put", "args": [lit(m.group(i)) for i in (1, 2, 3)]})
            continue
        m = re.match(r"CHECK\(arr\.size\(\) == (\d+)\);", s)
        if m:
            case["ops"].append({"op": "size", "expect": int(m.group(1))})
            continue
        m = re.match(r"CHECK\(arr\[(\d+)\]\.(start|end|delta_t) == (" + NUM + r")\);", s)
        if m:
            case["ops"].append({"op": "field", "index": int(m.group(1)), "field": m.group(2), "expect": lit(m.group(3))})
            continue
        m = re.match(r"auto v = DensityArray::combine\(\{\s*(" + NUM + r"),\s*(" + NUM + r"),\s*(" + NUM + r")\s*\},\s*\{\s*(" + NUM
                     + r"),\s*(" + NUM + r"),\s*(" + NUM + r")\s*\}\);", s)
        if m:
            case = {"section": section, "kind": "combine", "a": [lit(m.group(i)) for i in (1, 2, 3)],
                    "b": [lit(m.group(i)) for i in (4, 5, 6)], "checks": []}
            cases.append(case)
            continue
        m = re.match(r"CHECK\(v\[(\d+)\]\.(start|end|delta_t) == (" + NUM + r")\);", s)
        if m:
            case["checks"].append({"index": int(m.group(1)), "field": m.group(2), "expect": lit(m.group(3))})
            continue
        if s.startswith("CHECK(") or s.startswith("arr.") or ("DensityArray" in s and "using" not in s and "TEST_CASE" not in s):
            raise SystemExit(f"unparsed line: {s}")
    with open(out, "w") as f:
        json.dump({"source": "unittests/marching_steps.cpp:68-938", "cases": cases}, f, indent=1)
    print(f"{len(cases)} cases, {sum(len(c.get('ops', c.get('checks', []))) for c in cases)} checks -> {out}")


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/unittests/marching_steps.cpp"
    main(src, os.path.join(os.path.dirname(os.path.abspath(__file__)), "density_kats.json"))
