#!/usr/bin/env python3
"""Extract the reference's 3-D golden test vectors into JSON fixtures.

Reads the numeric arrays of

* ``ExaHyPE/exahype/tests/kernels/testdata/generic_euler_testdata.cpp``
  (``#ifdef Dim3`` block, lines 1348-6199)
* ``ExaHyPE/exahype/tests/kernels/testdata/limiter_testdata.cpp``
  (``#ifdef Dim3`` block)
* ``ExaHyPE/exahype/tests/kernels/testdata/elasticity_testdata.cpp``
  (``#ifdef Dim2`` block: the linear-kernel vectors of the 2-D elasticity
  test, order 4, 9 variables + 3 parameters)

and writes ``generic_euler_3d.json`` / ``limiter_3d.json`` /
``elasticity_2d.json`` next to this
script: ``{namespace: {array_name: [values...]}}``.  Only data is kept
(inputs and expected outputs); the test logic lives in ``tests/``.

Run once where the reference is available:

    python3 tests/golden/extract_reference_vectors.py /root/reference
"""
import json
import os
import re
import sys

ARRAY_START = re.compile(r"const\s+double\s+([A-Za-z0-9_]+)\s*\[([0-9*\s]*)\]\s*=?\s*\{")
NUMBER = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")


def dim_block(text, dim=3):
    start = text.index(f"#ifdef Dim{dim}")
    end = text.index("#endif", start)
    return text[start:end]


def parse(block):
    """Walk the block, tracking namespace nesting by brace depth."""
    out = {}
    ns_stack = []  # (name, depth at which its body opened)
    depth = 0
    pos = 0
    token = re.compile(r"namespace\s+([A-Za-z0-9_]+)\s*\{|const\s+double|\{|\}")
    while True:
        m = token.search(block, pos)
        if not m:
            break
        t = m.group(0)
        if t.startswith("namespace"):
            depth += 1
            ns_stack.append((m.group(1), depth))
            pos = m.end()
        elif t == "{":
            depth += 1
            pos = m.end()
        elif t == "}":
            if ns_stack and ns_stack[-1][1] == depth:
                ns_stack.pop()
            depth -= 1
            pos = m.end()
        else:
            a = ARRAY_START.match(block, m.start())
            if not a:
                pos = m.end()
                continue
            close = block.index("}", a.end())
            body = re.sub(r"//[^\n]*", "", block[a.end():close])
            vals = [float(v) for v in NUMBER.findall(body)]
            if a.group(2).strip() and vals:  # declared size, possibly a product such as 9 * 5
                size = 1
                for f in a.group(2).split("*"):
                    size *= int(f)
                assert len(vals) == size, (ns_stack, a.group(1), len(vals))
            key = ns_stack[-1][0] if ns_stack else ""
            out.setdefault(key, {})[a.group(1)] = vals
            pos = close + 1
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    tdir = os.path.join(ref, "ExaHyPE", "exahype", "tests", "kernels", "testdata")
    here = os.path.dirname(os.path.abspath(__file__))
    for src, dst, dim in (("generic_euler_testdata.cpp", "generic_euler_3d.json", 3),
                          ("limiter_testdata.cpp", "limiter_3d.json", 3),
                          ("elasticity_testdata.cpp", "elasticity_2d.json", 2)):
        with open(os.path.join(tdir, src)) as f:
            data = parse(dim_block(f.read(), dim))
        data["_source"] = {"file": "ExaHyPE/exahype/tests/kernels/testdata/" + src,
                           "block": f"#ifdef Dim{dim}"}
        with open(os.path.join(here, dst), "w") as f:
            json.dump(data, f, indent=0)
        print(dst, {k: {n: len(v) for n, v in d.items()} for k, d in data.items() if k != "_source"})


if __name__ == "__main__":
    main()
