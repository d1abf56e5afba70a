"""Extract the velocity-set tables from the reference sources (as text) into
tests/golden/velocity_sets.json.

Reads src/velocitySet/D3Q19.cuh, D3Q27.cuh and D3Q7.cuh of the reference
checkout: for each lattice, the direction index sets of every
`if constexpr (Q == a || Q == b ...) return +-1;` branch of cx/cy/cz, and
the weight of each direction from w<Q>() (the w_0/w_1/w_2/w_3 fractions).
The JSON holds only the resulting numbers.  Run in a container that has
the reference at /root/reference (the GPU machines do not need it).
"""
from __future__ import annotations

import json
import re
import sys
from fractions import Fraction
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "src" / "velocitySet"
OUT = Path(__file__).resolve().parent / "velocity_sets.json"


def _function_body(text: str, name: str) -> str:
    i = text.index(f" {name}(")
    j = text.index("{", i)
    depth = 0
    for k in range(j, len(text)):
        depth += {"{": 1, "}": -1}.get(text[k], 0)
        if depth == 0:
            return text[j:k + 1]
    raise ValueError(name)


def _components(text: str, axis: str, q: int) -> list[int]:
    body = _function_body(text, f"c{axis}")
    comp = [0] * q
    for cond, val in re.findall(r"if constexpr \(([^)]*)\)\s*\{\s*return (-?\d+);", body):
        for idx in re.findall(r"Q == (\d+)", cond):
            comp[int(idx)] = int(val)
    return comp


def _weights(text: str, q: int) -> list[str]:
    consts = {}
    for name, num, den in re.findall(
            r"(w_\d)\(\) noexcept\s*\{\s*return static_cast<scalar_t>\(static_cast<double>\((\d+)\) / static_cast<double>\((\d+)\)\);",
            text):
        consts[name] = Fraction(int(num), int(den))
    body = _function_body(text, "w")
    w = [None] * q
    # branches: Q == 0 / Q >= a && Q <= b / else
    for i in range(q):
        val = None
        for cond, name in re.findall(r"if(?: constexpr)? \(([^)]*)\)\s*\{\s*return (w_\d)\(\);", body):
            m_eq = re.fullmatch(r"Q == (\d+)", cond.strip())
            m_rng = re.fullmatch(r"Q >= (\d+) && Q <= (\d+)", cond.strip())
            if (m_eq and int(m_eq.group(1)) == i) or (m_rng and int(m_rng.group(1)) <= i <= int(m_rng.group(2))):
                val = consts[name]
                break
        if val is None:
            val = consts[re.findall(r"else\s*\{\s*return (w_\d)\(\);", body)[-1]]
        w[i] = str(val)
    return w


def main() -> None:
    out = {}
    for name, q in (("D3Q7", 7), ("D3Q19", 19), ("D3Q27", 27)):
        text = (REF / f"{name}.cuh").read_text()
        cx, cy, cz = (_components(text, a, q) for a in "xyz")
        out[name] = {"c": [[cx[i], cy[i], cz[i]] for i in range(q)], "w": _weights(text, q)}
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
